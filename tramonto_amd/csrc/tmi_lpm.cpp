// The reference's LinProbMgr C ABI (src/dft_basic_lin_prob_mgr_wrapper.h:44-92,
// src/dft_hardsphere_lin_prob_mgr_wrapper.h:43-58) and the fill entry point
// fill_resid_and_matrix_control (src/dft_fill_control.c:39), backed by the
// device fill and the matrix-free Schur solve.
//
// Call sequence served (src/dft_newton.c:131-196 and
// src/NOXLOCA_Tramonto_Group.cpp:176-313):
//   initializeproblemvalues -> fill_resid_and_matrix_control(x, iter, 0)
//   -> finalizeproblemvalues -> setupsolver -> solve -> getlhs
// The fill adds -F to the right-hand side (the reference's insertrhsvalue
// calls) and, for a full fill, records x as the Jacobian's linearisation
// point; solve() then runs GMRES on the Schur complement with the device JVP.
//
// Entries inserted through insert*matrixvalue* are kept (summed, as the
// reference's += semantics) for getmatrixvalue/writeMatrix only; solve()
// always uses the matrix-free Jacobian of the registered problem.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <vector>

#include "tmi_internal.hpp"

namespace tmi {
int schur_solve_export(Problem &P, const double *x, const double *F, double *dx, double rtol, int restart, int maxit,
                       int prec, int *iters, hipStream_t s);
Problem *problem_from_handle(void *h);

namespace {

struct LinProbMgr {
    int numUnks = 0;
    std::vector<int> indNonLocal, depNonLocal, density;
    bool a22diag = true;
    std::vector<int> rowmap, colmap;
    bool finalized = false;
    Problem *prob = nullptr;
    std::vector<double> rhs;   // [numUnks][Nowned]
    std::vector<double> lhs;
    std::map<std::pair<long long, long long>, double> inserted;
    double *d_x = nullptr, *d_F = nullptr, *d_dx = nullptr, *d_v = nullptr;
    bool have_x = false;
    double gmres_tol = 1e-8;
    int gmres_restart = 100, gmres_maxit = 500;
    int last_iters = 0;
};

LinProbMgr *g_active = nullptr;
Problem *g_active_problem = nullptr;

LinProbMgr *lpm(void *h)
{
    if (!h) throw std::invalid_argument("null LinProbMgr handle");
    return (LinProbMgr *)h;
}

void ensure_device(LinProbMgr &M)
{
    if (!M.prob) throw std::logic_error("no problem registered (tmi_set_active_problem)");
    if (M.d_x) return;
    size_t n = (size_t)M.prob->nunk * M.prob->N;
    TMI_HIP_CHECK(hipMalloc((void **)&M.d_x, n * sizeof(double)));
    TMI_HIP_CHECK(hipMalloc((void **)&M.d_F, n * sizeof(double)));
    TMI_HIP_CHECK(hipMalloc((void **)&M.d_dx, n * sizeof(double)));
    TMI_HIP_CHECK(hipMalloc((void **)&M.d_v, n * sizeof(double)));
}

void check_layout(const LinProbMgr &M)
{
    if (M.numUnks != M.prob->nunk_ext) throw std::logic_error("numUnks differs from the registered problem");
}

// host double** [nunk_ext][N] <-> contiguous device buffer in the internal layout; with attraction in
// A22 (no mean-field unknowns) the device buffer's mean-field rows are u * rho of its densities
// (linear: a direction, zero boundary values)
void to_device(Problem &P, double **x, double *d, bool linear)
{
    for (int u = 0; u < P.nunk_ext; ++u)
        TMI_HIP_CHECK(hipMemcpy(d + (size_t)u * P.N, x[u], sizeof(double) * P.N, hipMemcpyHostToDevice));
    if (P.ext_mf) launch_mf_conv(P, d + (size_t)P.iu_rho * P.N, linear, d + (size_t)P.iu_mf * P.N, nullptr);
}
void to_host(const Problem &P, const double *d, double **x)
{
    for (int u = 0; u < P.nunk_ext; ++u)
        TMI_HIP_CHECK(hipMemcpy(x[u], d + (size_t)u * P.N, sizeof(double) * P.N, hipMemcpyDeviceToHost));
}

template <typename F>
int guarded(F &&fn)
{
    try {
        return fn();
    } catch (const HipFailure &) {
        return TMI_EHIP;
    } catch (const std::logic_error &e) {
        set_error(e.what());
        return TMI_ESTATE;
    } catch (const std::exception &e) {
        set_error(e.what());
        return TMI_EINVAL;
    }
}

}  // namespace
}  // namespace tmi

using namespace tmi;

extern "C" {

int tmi_set_active_problem(void *problem_handle, void *linprobmgr_handle)
{
    return guarded([&] {
        Problem *P = problem_handle ? problem_from_handle(problem_handle) : nullptr;
        if (P && P->npores > 0)
            throw std::invalid_argument("a batched problem has no single-problem LinProbMgr counterpart");
        g_active_problem = P;
        g_active = (LinProbMgr *)linprobmgr_handle;
        if (g_active) g_active->prob = g_active_problem;
        return 0;
    });
}

// src/dft_fill_control.c:39.  resid_only_flag: 0 full fill (records the
// linearisation point), 1 residual only, 2/3 compute without inserting.
double fill_resid_and_matrix_control(double **x, int iter, int resid_only_flag)
{
    (void)iter;
    double l2 = -1.0;
    int rc = guarded([&] {
        if (!g_active_problem) throw std::logic_error("no active problem");
        Problem &P = *g_active_problem;
        LinProbMgr tmp;
        LinProbMgr &M = g_active ? *g_active : tmp;
        M.prob = &P;
        ensure_device(M);
        to_device(P, x, M.d_x, false);
        launch_residual(P, M.d_x, M.d_F, nullptr);
        const size_t n = (size_t)P.nunk_ext * P.N;   // the caller's rows
        std::vector<double> F(n);
        TMI_HIP_CHECK(hipMemcpy(F.data(), M.d_F, n * sizeof(double), hipMemcpyDeviceToHost));
        l2 = 0.0;
        for (double f : F) l2 += f * f;
        if (g_active && resid_only_flag != 2 && resid_only_flag != 3) {
            if (M.rhs.size() != n) M.rhs.assign(n, 0.0);
            for (size_t i = 0; i < n; ++i) M.rhs[i] += -F[i];
        }
        if (g_active && resid_only_flag == 0) M.have_x = true;
        if (!g_active) {
            (void)hipFree(M.d_x);
            (void)hipFree(M.d_F);
            (void)hipFree(M.d_dx);
            (void)hipFree(M.d_v);
        }
        return 0;
    });
    if (rc != 0) {
        std::fprintf(stderr, "fill_resid_and_matrix_control: %s\n", tmi_last_error());
        return -1.0;
    }
    return l2;
}

void *dft_hardsphere_lin_prob_mgr_create(int numUnks, void *ParameterList, void *comm)
{
    (void)ParameterList;
    (void)comm;
    auto M = new LinProbMgr();
    M->numUnks = numUnks;
    return M;
}

int dft_hardsphere_lin_prob_mgr_setindnonlocalequationids(void *h, int n, int *ids)
{
    return guarded([&] {
        lpm(h)->indNonLocal.assign(ids, ids + n);
        return 0;
    });
}
int dft_hardsphere_lin_prob_mgr_setdepnonlocalequationids(void *h, int n, int *ids)
{
    return guarded([&] {
        lpm(h)->depNonLocal.assign(ids, ids + n);
        return 0;
    });
}
int dft_hardsphere_lin_prob_mgr_setdensityequationids(void *h, int n, int *ids)
{
    return guarded([&] {
        lpm(h)->density.assign(ids, ids + n);
        return 0;
    });
}
int dft_hardsphere_lin_prob_mgr_seta22blockisdiagonal(void *h, int isA22Diagonal)
{
    return guarded([&] {
        lpm(h)->a22diag = isA22Diagonal != 0;
        return 0;
    });
}

int dft_linprobmgr_destruct(void *h)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        if (g_active == M) g_active = nullptr;
        (void)hipFree(M->d_x);
        (void)hipFree(M->d_F);
        (void)hipFree(M->d_dx);
        (void)hipFree(M->d_v);
        delete M;
        return 0;
    });
}

int dft_linprobmgr_setnodalrowmap(void *h, int n, int *gids)
{
    return guarded([&] {
        lpm(h)->rowmap.assign(gids, gids + n);
        return 0;
    });
}
int dft_linprobmgr_setnodalcolmap(void *h, int n, int *gids)
{
    return guarded([&] {
        lpm(h)->colmap.assign(gids, gids + n);
        return 0;
    });
}
int dft_linprobmgr_setcoarsenednodeslist(void *h, int n, int *gids)
{
    return guarded([&] {
        (void)gids;
        if (n > 0) throw std::domain_error("mesh coarsening is not supported");
        lpm(h);
        return 0;
    });
}

// 0 ok, 1 already finalised, -1 a set call is missing (dft_HardSphereLinProbMgr.cpp:71-90)
int dft_linprobmgr_finalizeblockstructure(void *h)
{
    LinProbMgr *M = (LinProbMgr *)h;
    if (!M) return -1;
    if (M->finalized) return 1;
    if (M->rowmap.empty() || M->colmap.empty() || M->indNonLocal.empty() || M->density.empty()) return -1;
    M->finalized = true;
    return 0;
}

int dft_linprobmgr_initializeproblemvalues(void *h)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        if (!M->finalized) throw std::logic_error("finalizeblockstructure not called");
        std::fill(M->rhs.begin(), M->rhs.end(), 0.0);
        M->inserted.clear();
        return 0;
    });
}

static size_t rhs_index(LinProbMgr *M, int iunk, int node)
{
    size_t nown = M->rowmap.size();
    if (iunk < 0 || iunk >= M->numUnks || node < 0 || (size_t)node >= nown) throw std::invalid_argument("row index");
    if (M->rhs.size() != (size_t)M->numUnks * nown) M->rhs.assign((size_t)M->numUnks * nown, 0.0);
    return (size_t)iunk * nown + node;
}

int dft_linprobmgr_insertrhsvalue(void *h, int iunk, int inode, double value)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        M->rhs[rhs_index(M, iunk, inode)] += value;
        return 0;
    });
}

int dft_linprobmgr_insertonematrixvalue(void *h, int iunk, int ownednode, int boxphysicsID, int boxnode, double value)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        long long nown = (long long)M->rowmap.size(), nbox = (long long)M->colmap.size();
        M->inserted[{(long long)iunk * nown + ownednode, (long long)boxphysicsID * nbox + boxnode}] += value;
        return 0;
    });
}

int dft_linprobmgr_insertmultinodematrixvalues(void *h, int iunk, int ownednode, int boxphysicsID, int *boxnodes,
                                               double *values, int numEntries)
{
    for (int k = 0; k < numEntries; ++k) {
        int rc = dft_linprobmgr_insertonematrixvalue(h, iunk, ownednode, boxphysicsID, boxnodes[k], values[k]);
        if (rc) return rc;
    }
    return 0;
}

int dft_linprobmgr_insertmultiphysicsmatrixvalues(void *h, int iunk, int ownednode, int *boxphysicsIDs, int boxnode,
                                                  double *values, int numEntries)
{
    for (int k = 0; k < numEntries; ++k) {
        int rc = dft_linprobmgr_insertonematrixvalue(h, iunk, ownednode, boxphysicsIDs[k], boxnode, values[k]);
        if (rc) return rc;
    }
    return 0;
}

int dft_linprobmgr_finalizeproblemvalues(void *h)
{
    return guarded([&] {
        lpm(h);
        return 0;
    });
}

int dft_linprobmgr_setrhs(void *h, double **b)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        size_t nown = M->rowmap.size();
        M->rhs.assign((size_t)M->numUnks * nown, 0.0);
        for (int u = 0; u < M->numUnks; ++u) std::memcpy(&M->rhs[(size_t)u * nown], b[u], sizeof(double) * nown);
        return 0;
    });
}

int dft_linprobmgr_getrhs(void *h, double **b)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        size_t nown = M->rowmap.size();
        if (M->rhs.size() != (size_t)M->numUnks * nown) M->rhs.assign((size_t)M->numUnks * nown, 0.0);
        for (int u = 0; u < M->numUnks; ++u) std::memcpy(b[u], &M->rhs[(size_t)u * nown], sizeof(double) * nown);
        return 0;
    });
}

int dft_linprobmgr_getlhs(void *h, double **x)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        size_t nbox = M->colmap.size();
        if (M->lhs.size() != (size_t)M->numUnks * nbox) throw std::logic_error("solve has not run");
        for (int u = 0; u < M->numUnks; ++u) std::memcpy(x[u], &M->lhs[(size_t)u * nbox], sizeof(double) * nbox);
        return 0;
    });
}

int dft_linprobmgr_setupsolver(void *h)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        if (!M->prob) throw std::logic_error("no problem registered");
        check_layout(*M);
        return 0;
    });
}

// J dx = rhs with the Jacobian at the last full fill
int dft_linprobmgr_solve(void *h)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        if (!M->have_x) throw std::logic_error("solve requires a full fill first");
        Problem &P = *M->prob;
        ensure_device(*M);
        const size_t n = (size_t)P.nunk_ext * P.N, ni = (size_t)P.nunk * P.N;
        std::vector<double> negb(ni, 0.0);   // internal mean-field rows (if any) stay 0: x_mf = u * rho holds
        for (size_t i = 0; i < n; ++i) negb[i] = -M->rhs[i];   // F such that rhs = -F
        TMI_HIP_CHECK(hipMemcpy(M->d_F, negb.data(), ni * sizeof(double), hipMemcpyHostToDevice));
        schur_solve_export(P, M->d_x, M->d_F, M->d_dx, M->gmres_tol, M->gmres_restart, M->gmres_maxit, TMI_PREC_FP64,
                           &M->last_iters, nullptr);
        M->lhs.assign(n, 0.0);
        TMI_HIP_CHECK(hipMemcpy(M->lhs.data(), M->d_dx, n * sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    });
}

// b = J x (dft_HardSphereLinProbMgr.cpp:430, ApplyGlobal)
int dft_linprobmgr_applymatrix(void *h, double **x, double **b)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        if (!M->have_x) throw std::logic_error("applymatrix requires a full fill first");
        Problem &P = *M->prob;
        to_device(P, x, M->d_v, true);
        launch_jvp<double>(P, M->d_x, M->d_v, M->d_dx, nullptr);
        TMI_HIP_CHECK(hipDeviceSynchronize());
        to_host(P, M->d_dx, b);
        return 0;
    });
}

// one GPU holds the whole domain: owned == box
int dft_linprobmgr_importr2c(void *h, double **xOwned, double **xBox)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        size_t n = M->colmap.size();
        if (M->rowmap.size() != n) throw std::domain_error("owned and box maps differ");
        for (int u = 0; u < M->numUnks; ++u)
            if (xBox[u] != xOwned[u]) std::memcpy(xBox[u], xOwned[u], sizeof(double) * n);
        return 0;
    });
}

int dft_linprobmgr_importnodalr2c(void *h, double *xOwned, double *xBox)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        if (xBox != xOwned) std::memcpy(xBox, xOwned, sizeof(double) * M->colmap.size());
        return 0;
    });
}

int dft_linprobmgr_importsingleunknownr2c(void *h, double *xOwned, double *xBox)
{
    return dft_linprobmgr_importnodalr2c(h, xOwned, xBox);
}

// J[row, col] at the last full fill: from inserted entries if any were
// inserted, else one device JVP with a unit vector (tests only).
double dft_linprobmgr_getmatrixvalue(void *h, int iunk, int ownednode, int junk, int boxnode)
{
    double val = 0.0;
    int rc = guarded([&] {
        LinProbMgr *M = lpm(h);
        long long nown = (long long)M->rowmap.size(), nbox = (long long)M->colmap.size();
        if (!M->inserted.empty()) {
            auto it = M->inserted.find({(long long)iunk * nown + ownednode, (long long)junk * nbox + boxnode});
            val = (it == M->inserted.end()) ? 0.0 : it->second;
            return 0;
        }
        if (!M->have_x) throw std::logic_error("getmatrixvalue requires a full fill first");
        Problem &P = *M->prob;
        const size_t n = (size_t)P.nunk * P.N;
        if (junk < 0 || junk >= P.nunk_ext || iunk < 0 || iunk >= P.nunk_ext) throw std::invalid_argument("unknown index");
        TMI_HIP_CHECK(hipMemset(M->d_v, 0, n * sizeof(double)));
        double one = 1.0;
        TMI_HIP_CHECK(hipMemcpy(M->d_v + (size_t)junk * P.N + boxnode, &one, sizeof(double), hipMemcpyHostToDevice));
        if (P.ext_mf) launch_mf_conv(P, M->d_v + (size_t)P.iu_rho * P.N, true, M->d_v + (size_t)P.iu_mf * P.N, nullptr);
        launch_jvp<double>(P, M->d_x, M->d_v, M->d_dx, nullptr);
        TMI_HIP_CHECK(hipMemcpy(&val, M->d_dx + (size_t)iunk * P.N + ownednode, sizeof(double), hipMemcpyDeviceToHost));
        return 0;
    });
    return rc == 0 ? val : 0.0;
}

int dft_linprobmgr_writeMatrix(void *h, char *filename, char *matrixName, char *matrixDescription)
{
    return guarded([&] {
        LinProbMgr *M = lpm(h);
        if (!M->have_x) throw std::logic_error("writeMatrix requires a full fill first");
        Problem &P = *M->prob;
        const size_t n = (size_t)P.nunk_ext * P.N, ni = (size_t)P.nunk * P.N;
        if (n > 20000) throw std::domain_error("writeMatrix is limited to small problems");
        FILE *f = std::fopen(filename, "w");
        if (!f) throw std::invalid_argument("cannot open matrix file");
        std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%% %s: %s\n", matrixName ? matrixName : "",
                     matrixDescription ? matrixDescription : "");
        std::vector<double> col(n);
        std::vector<std::vector<std::pair<size_t, double>>> cols(n);
        size_t nnz = 0;
        for (size_t j = 0; j < n; ++j) {
            TMI_HIP_CHECK(hipMemset(M->d_v, 0, ni * sizeof(double)));
            double one = 1.0;
            TMI_HIP_CHECK(hipMemcpy(M->d_v + j, &one, sizeof(double), hipMemcpyHostToDevice));
            if (P.ext_mf)
                launch_mf_conv(P, M->d_v + (size_t)P.iu_rho * P.N, true, M->d_v + (size_t)P.iu_mf * P.N, nullptr);
            launch_jvp<double>(P, M->d_x, M->d_v, M->d_dx, nullptr);
            TMI_HIP_CHECK(hipMemcpy(col.data(), M->d_dx, n * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i)
                if (col[i] != 0.0) {
                    cols[j].push_back({i, col[i]});
                    ++nnz;
                }
        }
        std::fprintf(f, "%zu %zu %zu\n", n, n, nnz);
        for (size_t j = 0; j < n; ++j)
            for (auto &e : cols[j]) std::fprintf(f, "%zu %zu %.17g\n", e.first + 1, j + 1, e.second);
        std::fclose(f);
        return 0;
    });
}

}  // extern "C"
