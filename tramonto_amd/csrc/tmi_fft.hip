// The mean-field (attraction) sums of 2D/3D problems as FFT convolutions.
//
// The mean-field rows are a correlation of the densities with the attraction
// stencils, F_MF,c(i) = sum_c' sum_k u_cc'(k) rho_c'(i + k) - x_MF,c(i)
// (load_mean_field, src/dft_fill_mf.c:38-63, through
// resid_and_Jac_sten_fill_sum_Ncomp, src/dft_fill_CLSmf.c:38-171).  With tens
// of thousands of stencil points per node the direct form is bound by FP64
// FMAs; as a circular convolution on an FFT grid it costs a few HBM-bound
// passes instead.
//
// FFT grid, per dimension: a periodic axis whose node count factors into 2,
// 3 and 5 is used as is (a circular convolution is the periodic image sum);
// any other axis uses the halo-padded copy of the density (the halo holds the
// boundary values offset_to_node_box gives) extended to a 2^a 3^b 5^c length.
// Interior outputs of a padded axis never wrap, because the halo is the
// stencil reach.  The transforms are mixed-radix Stockham passes in LDS, one
// launch per axis; the kernel spectra (flipped stencils, scaled by 1/M) are
// computed once at set-up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "tmi_device.hpp"
#include "tmi_internal.hpp"

namespace tmi {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
    T x, y;
};

struct FftPlan {
    int M[3] = {1, 1, 1};     // FFT grid
    int off[3] = {0, 0, 0};   // FFT-grid coordinate of node 0
    int src[3] = {0, 0, 0};   // density-copy coordinate of FFT-grid position 0
    int lim[3] = {1, 1, 1};   // FFT-grid positions below lim map onto the density copy
    int rc[3][4] = {};          // Stockham stages per axis: counts of radix 4, 2, 3 and 5, run in that order
    long long n = 1;            // grid points M0 M1 M2
    long long S1 = 1, S2 = 1;   // storage pitches of y rows and z planes (row_pitch; one row of padding
                                // per plane), so that no stride of the strided passes is a power of two
    long long ns = 1;           // storage elements per field
    Cx<double> *kd = nullptr;   // [C*C][n] spectra, index icomp + C*jcomp as the stencils
    Cx<float> *kf = nullptr;
    Cx<double> *twd[3] = {nullptr, nullptr, nullptr};   // exp(-2 pi i j / M_d), j < M_d
    Cx<float> *twf[3] = {nullptr, nullptr, nullptr};
    void *buf = nullptr;        // 2 C n complex doubles: transforms of the inputs, then of the outputs
};

namespace {

using namespace dev;

constexpr int kFftBlock = 256;
constexpr size_t kFftLds = 48 * 1024;
constexpr int kFftMaxLen = 768;   // longest FFT axis: its twiddles and 2 lines of complex doubles fit kFftLds

// Stockham stages per axis in the order they run: radix 4, then 2, then 3, then 5
struct RadixCounts {
    int n4, n2, n3, n5;
};

template <typename T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ inline Cx<T> cadd(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ inline Cx<T> csub(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ inline Cx<T> cscale(Cx<T> a, T s)
{
    return Cx<T>{a.x * s, a.y * s};
}
// i * a * SIGN
template <int SIGN, typename T>
__device__ inline Cx<T> cis(Cx<T> a)
{
    if constexpr (SIGN > 0)
        return Cx<T>{-a.y, a.x};
    else
        return Cx<T>{a.y, -a.x};
}

// X_p = sum_q a_q exp(SIGN 2 pi i p q / R), in place, in registers
template <int R, int SIGN, typename T>
__device__ inline void dft_small(Cx<T> *a)
{
    if constexpr (R == 2) {
        const Cx<T> t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if constexpr (R == 4) {
        const Cx<T> t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        const Cx<T> t2 = cadd(a[1], a[3]), t3 = cis<SIGN>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    } else if constexpr (R == 3) {
        const T h = (T)0.86602540378443864676;   // sqrt(3)/2
        const Cx<T> s = cadd(a[1], a[2]), d = cis<SIGN>(cscale(csub(a[1], a[2]), h));
        const Cx<T> m = csub(a[0], cscale(s, (T)0.5));
        a[0] = cadd(a[0], s);
        a[1] = cadd(m, d);
        a[2] = csub(m, d);
    } else {
        static_assert(R == 5, "radix 2, 3, 4 or 5");
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        const Cx<T> p1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
        const Cx<T> p2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
        const Cx<T> b1 = cadd(a[0], cadd(cscale(p1, c1), cscale(p2, c2)));
        const Cx<T> b2 = cadd(a[0], cadd(cscale(p1, c2), cscale(p2, c1)));
        const Cx<T> e1 = cis<SIGN>(cadd(cscale(d1, s1), cscale(d2, s2)));
        const Cx<T> e2 = cis<SIGN>(csub(cscale(d1, s2), cscale(d2, s1)));
        a[0] = cadd(a[0], cadd(p1, p2));
        a[1] = cadd(b1, e1);
        a[4] = csub(b1, e1);
        a[2] = cadd(b2, e2);
        a[3] = csub(b2, e2);
    }
}

// q = x / d for 0 <= x < 2^22 through a float reciprocal, corrected to exact
__device__ inline int small_div(int x, int d, float inv)
{
    int q = (int)((float)x * inv);
    if ((q + 1) * d <= x) ++q;
    if (q * d > x) --q;
    return q;
}

// One radix-R Stockham stage over nl lines of length M held in LDS (row pitch
// ld): src -> dst, with Ns the length of the sub-transforms already done.
// The butterfly is a compile-time radix, so its operands stay in registers.
template <int R, int SIGN, typename T>
__device__ inline void stockham_stage(const Cx<T> *__restrict__ src, Cx<T> *__restrict__ dst, int ld, int nl, int M,
                                      int Ns, const Cx<T> *__restrict__ tw)
{
    const int nb = M / R;
    const int tstep = M / (Ns * R);   // twiddle exp(SIGN 2 pi i r k / (Ns R)) = tw[r k tstep]
    const float inv_nb = 1.0f / (float)nb, inv_ns = 1.0f / (float)Ns;
    for (int t = threadIdx.x; t < nl * nb; t += kFftBlock) {
        const int i = small_div(t, nb, inv_nb), jj = t - i * nb;
        const int q = small_div(jj, Ns, inv_ns), k = jj - q * Ns;
        const Cx<T> *a = src + (size_t)i * ld + jj;
        Cx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = a[r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                Cx<T> w = tw[r * k * tstep];
                if constexpr (SIGN > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<R, SIGN>(v);
        Cx<T> *b = dst + (size_t)i * ld + q * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) b[r * Ns] = v[r];
    }
}

// One axis of a 3D transform, in place.  A workgroup takes lpb lines of the
// axis (consecutive lines along x for the strided axes, so that the loads
// are coalesced), runs every Stockham stage in LDS and writes back.
template <typename T, int SIGN>
__global__ __launch_bounds__(kFftBlock) void k_fft_axis(Cx<T> *__restrict__ data, int M0, int M1, int M2, long long S1,
                                                        long long S2, int axis, RadixCounts rc, int lpb,
                                                        const Cx<T> *__restrict__ twtab)
{
    extern __shared__ char smem_raw[];
    const int M = axis == 0 ? M0 : (axis == 1 ? M1 : M2);
    const int ld = M + 1;
    Cx<T> *tw = reinterpret_cast<Cx<T> *>(smem_raw);   // [M] twiddles exp(-2 pi i j / M)
    Cx<T> *bufA = tw + M;
    Cx<T> *bufB = bufA + (size_t)lpb * ld;
    for (int j = threadIdx.x; j < M; j += kFftBlock) tw[j] = twtab[j];
    // line geometry
    long long base0 = 0;
    long long step = 1;   // between elements of a line (strided axes)
    int nl;               // lines in this group
    int l0 = 0;           // first line of the group (axis 0: line = y + M1 z)
    if (axis == 0) {
        const int nlines = M1 * M2;
        l0 = blockIdx.x * lpb;
        nl = (nlines - l0) < lpb ? (nlines - l0) : lpb;
    } else {
        const int xg = (M0 + lpb - 1) / lpb;
        const int other = blockIdx.x / xg, x0 = (blockIdx.x % xg) * lpb;
        nl = (M0 - x0) < lpb ? (M0 - x0) : lpb;
        if (axis == 1) {
            base0 = x0 + S2 * other;
            step = S1;
        } else {
            base0 = x0 + S1 * other;
            step = S2;
        }
    }
    const float inv_m = 1.0f / (float)M, inv_nl = 1.0f / (float)nl, inv_m1 = 1.0f / (float)M1;
    // storage index of element j of line i of an axis-0 group
    auto row0 = [&](int i) {
        const int l = l0 + i, z = small_div(l, M1, inv_m1);
        return (long long)(l - z * M1) * S1 + (long long)z * S2;
    };
    // load
    if (axis == 0) {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int i = small_div(e, M, inv_m), j = e - i * M;
            bufA[i * ld + j] = data[row0(i) + j];
        }
    } else {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int j = small_div(e, nl, inv_nl), i = e - j * nl;
            bufA[i * ld + j] = data[base0 + i + (long long)j * step];
        }
    }
    __syncthreads();
    int Ns = 1;
    auto run = [&](auto radix, int count) {
        constexpr int R = decltype(radix)::value;
        for (int s = 0; s < count; ++s) {
            stockham_stage<R, SIGN>(bufA, bufB, ld, nl, M, Ns, tw);
            __syncthreads();
            Cx<T> *tmp = bufA;
            bufA = bufB;
            bufB = tmp;
            Ns *= R;
        }
    };
    run(std::integral_constant<int, 4>{}, rc.n4);
    run(std::integral_constant<int, 2>{}, rc.n2);
    run(std::integral_constant<int, 3>{}, rc.n3);
    run(std::integral_constant<int, 5>{}, rc.n5);
    // store
    if (axis == 0) {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int i = small_div(e, M, inv_m), j = e - i * M;
            data[row0(i) + j] = bufA[i * ld + j];
        }
    } else {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int j = small_div(e, nl, inv_nl), i = e - j * nl;
            data[base0 + i + (long long)j * step] = bufA[i * ld + j];
        }
    }
}

struct FftGeo {
    int M[3], src[3], lim[3], off[3];
    long long n, S1, S2, ns;
};

// inputs: R_c'(m) = density copy at the FFT-grid position (0 beyond the copy)
template <typename T>
__global__ void k_fft_gather(FftGeo f, Geo g, const T *__restrict__ rho_pad, Cx<T> *__restrict__ R)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m >= f.n) return;
    const int m0 = (int)(m % f.M[0]);
    const long long r = m / f.M[0];
    const int m1 = (int)(r % f.M[1]), m2 = (int)(r / f.M[1]);
    const bool in = m0 < f.lim[0] && m1 < f.lim[1] && m2 < f.lim[2];
    const long long p = in ? (long long)(m0 + f.src[0]) + (long long)(m1 + f.src[1]) * g.P[0] +
                                 (long long)(m2 + f.src[2]) * g.P[0] * g.P[1]
                           : 0;
    const long long q = m0 + m1 * f.S1 + m2 * f.S2;
    for (int c = 0; c < g.C; ++c) R[(size_t)c * f.ns + q] = Cx<T>{in ? rho_pad[(size_t)c * g.np + p] : (T)0, (T)0};
}

// O_c = sum_c' K[c + C c'] R_c'
template <typename T>
__global__ void k_fft_mul(int C, long long n, const Cx<T> *__restrict__ K, const Cx<T> *__restrict__ R,
                          Cx<T> *__restrict__ O)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m >= n) return;
    for (int c = 0; c < C; ++c) {
        Cx<T> acc{(T)0, (T)0};
        for (int cj = 0; cj < C; ++cj) acc = cadd(acc, cmul(K[(size_t)(c + C * cj) * n + m], R[(size_t)cj * n + m]));
        O[(size_t)c * n + m] = acc;
    }
}

// out_mf[c][i] = Re O_c(node i) - in_mf - sub
template <typename T>
__global__ void k_fft_extract(FftGeo f, Geo g, const Cx<T> *__restrict__ O, const double *__restrict__ in_mf,
                              const double *__restrict__ sub, double *__restrict__ out_mf)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= g.N) return;
    const int x = i % g.nx[0];
    const int r = i / g.nx[0];
    const int y = r % g.nx[1], z = r / g.nx[1];
    const long long m = (long long)(x + f.off[0]) + (long long)(y + f.off[1]) * f.S1 + (long long)(z + f.off[2]) * f.S2;
    const size_t row = (size_t)c * g.N + i;
    double v = (double)O[(size_t)c * f.ns + m].x;
    if (in_mf) v -= in_mf[row];
    if (sub) v -= sub[row];
    out_mf[row] = v;
}

template <typename T>
__global__ void k_cast_spectrum(const Cx<double> *__restrict__ a, Cx<T> *__restrict__ b, long long n)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m < n) b[m] = Cx<T>{(T)a[m].x, (T)a[m].y};
}

__global__ void k_scale_cx(Cx<double> *__restrict__ a, double s, long long n)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m < n) a[m] = Cx<double>{a[m].x * s, a[m].y * s};
}

// Row pitch of the FFT storage: a multiple of 16 elements (rows start on 128-byte lines in complex
// float and complex double), at least one 16-element pad past the row, and not a power of two.
long long row_pitch(int M0)
{
    long long s = ((M0 + 15) / 16) * 16 + 16;
    if ((s & (s - 1)) == 0) s += 16;
    return s;
}

int nice_size(int n)
{
    for (int m = std::max(n, 1);; ++m) {
        int r = m;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

// M = 4^c[0] 2^c[1] 3^c[2] 5^c[3] with c[1] <= 1
void factor_radices(int M, int *c)
{
    int r = M;
    const int radix[4] = {4, 2, 3, 5};
    for (int i = 0; i < 4; ++i) {
        c[i] = 0;
        while (r % radix[i] == 0) {
            ++c[i];
            r /= radix[i];
        }
    }
    if (r != 1) throw std::logic_error("FFT length must factor into 2, 3 and 5");
}

template <typename T>
inline const Cx<T> *twiddles(const FftPlan &F, int axis);
template <>
inline const Cx<double> *twiddles<double>(const FftPlan &F, int axis)
{
    return F.twd[axis];
}
template <>
inline const Cx<float> *twiddles<float>(const FftPlan &F, int axis)
{
    return F.twf[axis];
}

// Lines per workgroup: as many as the LDS budget allows, but few enough that
// the pass has about kFftGroups workgroups (the whole chip busy).
constexpr long long kFftGroups = 1024;

template <typename T>
void fft3d(const FftPlan &F, Cx<T> *data, int sign, hipStream_t s)
{
    for (int axis = 0; axis < 3; ++axis) {
        const int M = F.M[axis];
        if (M == 1) continue;
        const long long lines = F.n / M;
        const long long lds_max = (long long)((kFftLds / sizeof(Cx<T>) - M) / (2 * (M + 1)));
        long long lpb = std::max<long long>(1, std::min<long long>(lds_max, lines / kFftGroups));
        lpb = std::min<long long>(lpb, 64);
        if (axis > 0) {   // strided axes: whole 128-byte lines of consecutive x when the budget allows
            const long long line_elems = 128 / (long long)sizeof(Cx<T>);
            if (lds_max >= line_elems) lpb = std::max(line_elems, (lpb / line_elems) * line_elems);
        }
        long long groups;
        if (axis == 0)
            groups = (lines + lpb - 1) / lpb;
        else
            groups = (long long)((F.M[0] + lpb - 1) / lpb) * (axis == 1 ? F.M[2] : F.M[1]);
        const RadixCounts rc{F.rc[axis][0], F.rc[axis][1], F.rc[axis][2], F.rc[axis][3]};
        const size_t lds = sizeof(Cx<T>) * ((size_t)M + 2 * (size_t)lpb * (M + 1));
        if (sign < 0)
            hipLaunchKernelGGL((k_fft_axis<T, -1>), dim3((unsigned)groups), dim3(kFftBlock), lds, s, data, F.M[0], F.M[1],
                               F.M[2], F.S1, F.S2, axis, rc, (int)lpb, twiddles<T>(F, axis));
        else
            hipLaunchKernelGGL((k_fft_axis<T, 1>), dim3((unsigned)groups), dim3(kFftBlock), lds, s, data, F.M[0], F.M[1],
                               F.M[2], F.S1, F.S2, axis, rc, (int)lpb, twiddles<T>(F, axis));
    }
}

FftGeo fft_geo(const FftPlan &F)
{
    FftGeo f;
    for (int d = 0; d < 3; ++d) {
        f.M[d] = F.M[d];
        f.src[d] = F.src[d];
        f.lim[d] = F.lim[d];
        f.off[d] = F.off[d];
    }
    f.n = F.n;
    f.S1 = F.S1;
    f.S2 = F.S2;
    f.ns = F.ns;
    return f;
}

}  // namespace

// Plan, spectra and buffers of the FFT mean-field path (P.fft stays null when an
// axis would be longer than kFftMaxLen).
void fft_setup(Problem &P)
{
    const int nd = P.p.ndim, C = P.p.ncomp;
    auto *F = new FftPlan();
    try {
        for (int d = 0; d < nd; ++d) {
            const int n = P.p.nx[d], h = P.pad_u.h[d];
            const bool per = P.p.bc[d][0] == TMI_PERIODIC;
            if (per && nice_size(n) == n) {
                F->M[d] = n;
                F->off[d] = 0;
                F->src[d] = h;
                F->lim[d] = n;
            } else {
                F->M[d] = nice_size(n + 2 * h);
                F->off[d] = h;
                F->src[d] = 0;
                F->lim[d] = n + 2 * h;
            }
            factor_radices(F->M[d], F->rc[d]);
            if (F->M[d] > kFftMaxLen) {   // one line must fit the LDS budget: keep the direct sums
                delete F;
                return;
            }
        }
        F->n = (long long)F->M[0] * F->M[1] * F->M[2];
        F->S1 = F->M[1] > 1 ? row_pitch(F->M[0]) : F->M[0];
        F->S2 = F->S1 * (F->M[1] + (F->M[2] > 1 ? 1 : 0));
        F->ns = F->S2 * F->M[2];
        const size_t n = (size_t)F->ns;
        for (int d = 0; d < 3; ++d) {
            const int M = F->M[d];
            std::vector<Cx<double>> td(M);
            std::vector<Cx<float>> tf(M);
            for (int j = 0; j < M; ++j) {
                const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)M;
                td[j] = Cx<double>{(double)cosl(a), (double)sinl(a)};
                tf[j] = Cx<float>{(float)td[j].x, (float)td[j].y};
            }
            TMI_HIP_CHECK(hipMalloc((void **)&F->twd[d], M * sizeof(Cx<double>)));
            TMI_HIP_CHECK(hipMalloc((void **)&F->twf[d], M * sizeof(Cx<float>)));
            TMI_HIP_CHECK(hipMemcpy(F->twd[d], td.data(), M * sizeof(Cx<double>), hipMemcpyHostToDevice));
            TMI_HIP_CHECK(hipMemcpy(F->twf[d], tf.data(), M * sizeof(Cx<float>), hipMemcpyHostToDevice));
        }
        TMI_HIP_CHECK(hipMalloc(&F->buf, 2 * (size_t)C * n * sizeof(Cx<double>)));
        TMI_HIP_CHECK(hipMalloc((void **)&F->kd, (size_t)C * C * n * sizeof(Cx<double>)));
        TMI_HIP_CHECK(hipMalloc((void **)&F->kf, (size_t)C * C * n * sizeof(Cx<float>)));
        // flipped stencil g(-k mod M) = u(k) on the FFT grid, transformed, scaled by 1/n
        std::vector<Cx<double>> host(n);
        for (int q = 0; q < C * C; ++q) {
            std::fill(host.begin(), host.end(), Cx<double>{0.0, 0.0});
            const HostStencil &hs = P.hs[TMI_UATT][q];
            const long long pitch[3] = {1, F->S1, F->S2};
            for (int k = 0; k < hs.L; ++k) {
                long long idx = 0;
                for (int d = 0; d < 3; ++d) {
                    const int o = (d < nd) ? hs.off[3 * (size_t)k + d] : 0;
                    int j = (-o) % F->M[d];
                    if (j < 0) j += F->M[d];
                    idx += (long long)j * pitch[d];
                }
                host[(size_t)idx].x += hs.w[k];
            }
            Cx<double> *kq = F->kd + (size_t)q * n;
            TMI_HIP_CHECK(hipMemcpy(kq, host.data(), n * sizeof(Cx<double>), hipMemcpyHostToDevice));
            fft3d<double>(*F, kq, -1, 0);
            hipLaunchKernelGGL(k_scale_cx, dim3(grid_for((long long)n)), dim3(kBlock), 0, 0, kq, 1.0 / (double)F->n,
                               (long long)n);
            hipLaunchKernelGGL(k_cast_spectrum<float>, dim3(grid_for((long long)n)), dim3(kBlock), 0, 0, kq,
                               F->kf + (size_t)q * n, (long long)n);
        }
        TMI_HIP_CHECK(hipGetLastError());
        TMI_HIP_CHECK(hipDeviceSynchronize());
    } catch (...) {
        (void)hipFree(F->buf);
        (void)hipFree(F->kd);
        (void)hipFree(F->kf);
        for (int d = 0; d < 3; ++d) {
            (void)hipFree(F->twd[d]);
            (void)hipFree(F->twf[d]);
        }
        delete F;
        throw;
    }
    P.fft = F;
}

void fft_free(Problem &P)
{
    if (!P.fft) return;
    (void)hipFree(P.fft->buf);
    (void)hipFree(P.fft->kd);
    (void)hipFree(P.fft->kf);
    for (int d = 0; d < 3; ++d) {
        (void)hipFree(P.fft->twd[d]);
        (void)hipFree(P.fft->twf[d]);
    }
    delete P.fft;
    P.fft = nullptr;
}

// mean-field rows from the halo-padded densities in workspace<T>(P).rho_pad
template <typename T>
void mf_rows_fft(Problem &P, const T *rho_pad, const double *in_mf, const double *sub, double *out_mf, hipStream_t s)
{
    const FftPlan &F = *P.fft;
    const int C = P.p.ncomp;
    const size_t n = (size_t)F.ns;
    Cx<T> *R = reinterpret_cast<Cx<T> *>(F.buf);
    Cx<T> *O = R + (size_t)C * n;
    const Cx<T> *K = sizeof(T) == 8 ? reinterpret_cast<const Cx<T> *>(F.kd) : reinterpret_cast<const Cx<T> *>(F.kf);
    const Geo gu = make_geo(P, P.pad_u);
    const FftGeo fg = fft_geo(F);
    hipLaunchKernelGGL(k_fft_gather<T>, dim3(grid_for(F.n)), dim3(kBlock), 0, s, fg, gu, rho_pad, R);
    for (int c = 0; c < C; ++c) fft3d<T>(F, R + (size_t)c * n, -1, s);
    hipLaunchKernelGGL(k_fft_mul<T>, dim3(grid_for(F.ns)), dim3(kBlock), 0, s, C, F.ns, K, R, O);
    for (int c = 0; c < C; ++c) fft3d<T>(F, O + (size_t)c * n, +1, s);
    hipLaunchKernelGGL(k_fft_extract<T>, dim3(grid_for(P.N), C), dim3(kBlock), 0, s, fg, gu, O, in_mf, sub, out_mf);
    TMI_HIP_CHECK(hipGetLastError());
}
template void mf_rows_fft<double>(Problem &, const double *, const double *, const double *, double *, hipStream_t);
template void mf_rows_fft<float>(Problem &, const float *, const double *, const double *, double *, hipStream_t);

}  // namespace tmi
