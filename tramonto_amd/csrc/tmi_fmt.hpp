// Pointwise fundamental-measure free-energy derivatives, shared by host and
// device code.  Densities are in "n order": n0 n1 n2 n3 nv1[D] nv2[D].
//
// The four functionals are written as
//   Phi = -n0 ln(1-n3) + (n1 n2 - nv1.nv2)/(1-n3) + q(n2, nv2) f(n3) / (36 pi)
// with
//   FMT1 (Rosenfeld, src/dft_physics_FMT1.c:39-57):   q = n2^3 - 3 n2 nv2.nv2,   f = 1.5/(1-n3)^2
//   FMT2 (src/dft_physics_FMT2.c:40-58):              q = (n2 - nv2.nv2/n2)^3,   f = 1.5/(1-n3)^2
//   FMT3 (White Bear, src/dft_physics_FMT3.c:39-59):  q = n2^3 - 3 n2 nv2.nv2,   f = [n3 + (1-n3)^2 ln(1-n3)] / (n3^2 (1-n3)^2)
//   FMT4 (Oleksy-Hansen, src/dft_physics_FMT4.c:38-58): q = (n2 - nv2.nv2/n2)^3, f as White Bear.
// The first derivatives reproduce FMTk_1stderiv (src/dft_physics_FMT1.c:59-75, FMT2.c:60-100,
// FMT3.c:61-96, FMT4.c:60-96), including White Bear's and FMT4's n3 <= 1e-10 cut-off and FMT2's
// n2 <= 1e-15 convention (q, its derivatives taken at nv2 = 0), which FMT4 shares here (the reference
// divides by n2^2 there).  The Jacobian-vector product uses the exact Hessian of Phi; contracted with a
// stencil kernel vector it equals the reference's d2phi_drb2_{delta,theta}_rb values
// (src/dft_physics_FMT1.c:79-168, FMT3.c:101-266, and the theta parts of FMT2/FMT4), which are zero
// below the cut-off.  The reference's FMT2 (2D/3D) and FMT4 delta Hessians differ from the exact
// derivative in their nv2 columns; the product here is the exact one.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace tmi {

struct FmtScalars {
    double a;    // 1/(1-n3)
    double f0, f1, f2;   // f(n3), f'(n3), f''(n3)
    double lnA;  // ln(1/(1-n3))
    bool cut;    // White Bear below the n3 cut-off
};

// Rosenfeld's f(n3) for FMT1 and FMT2 (Type_func 0, 1), White Bear's for FMT3 and FMT4 (2, 3)
__host__ __device__ __attribute__((always_inline)) inline bool fmt_rosenfeld_f(int func) { return func <= 1; }
// q = (n2 - nv2.nv2/n2)^3 for FMT2 and FMT4, n2^3 - 3 n2 nv2.nv2 for FMT1 and FMT3
__host__ __device__ __attribute__((always_inline)) inline bool fmt_cubic_q(int func) { return func == 1 || func == 3; }

__host__ __device__ __attribute__((always_inline)) inline FmtScalars fmt_scalars(int func, double n3)
{
    FmtScalars s;
    const double one_m = 1.0 - n3;
    s.a = 1.0 / one_m;
    s.lnA = log(s.a);
    s.cut = false;
    const double a = s.a, a2 = a * a, a3 = a2 * a, a4 = a3 * a;
    if (fmt_rosenfeld_f(func)) {
        s.f0 = 1.5 * a2;
        s.f1 = 3.0 * a3;
        s.f2 = 9.0 * a4;
        return s;
    }
    if (!(n3 > 1.e-10)) {
        s.cut = true;
        s.f0 = s.f1 = s.f2 = 0.0;
        return s;
    }
    const double L = log(one_m);
    const double g = n3 + one_m * one_m * L;   // n3 + (1-n3)^2 ln(1-n3)
    const double g1 = n3 - 2.0 * one_m * L;    // g'
    const double g2 = 3.0 + 2.0 * L;           // g''
    const double i2 = 1.0 / (n3 * n3), i3 = i2 / n3, i4 = i3 / n3;
    s.f0 = g * a2 * i2;
    s.f1 = g1 * a2 * i2 + 2.0 * g * a3 * i2 - 2.0 * g * a2 * i3;
    s.f2 = g2 * a2 * i2 + 4.0 * g1 * a3 * i2 - 4.0 * g1 * a2 * i3 + 6.0 * g * a4 * i2 - 8.0 * g * a3 * i3 +
           6.0 * g * a2 * i4;
    return s;
}

// q(n2, nv2) and its derivatives, with D = nv2.nv2:
//   dq/dn2 = q2, dq/dnv2_d = qv nv2_d, d2q/dn2^2 = q22, d2q/dn2 dnv2_d = q2v nv2_d,
//   d2q/dnv2_d dnv2_e = qvvo nv2_d nv2_e + qvvd delta_de
struct FmtQ {
    double q, q2, qv, q22, q2v, qvvo, qvvd;
};

__host__ __device__ __attribute__((always_inline)) inline FmtQ fmt_q(int func, double n2, double D)
{
    FmtQ r;
    if (!fmt_cubic_q(func)) {   // n2^3 - 3 n2 D
        r.q = n2 * n2 * n2 - 3.0 * n2 * D;
        r.q2 = 3.0 * n2 * n2 - 3.0 * D;
        r.qv = -6.0 * n2;
        r.q22 = 6.0 * n2;
        r.q2v = -6.0;
        r.qvvo = 0.0;
        r.qvvd = -6.0 * n2;
        return r;
    }
    // alpha^3, alpha = n2 - D/n2, beta = d alpha/dn2 = 1 + D/n2^2; for n2 <= 1e-15 taken at nv2 = 0
    const bool small = !(n2 > 1.e-15);
    const double in2 = small ? 0.0 : 1.0 / n2;
    const double Dn = small ? 0.0 : D;
    const double al = n2 - Dn * in2, be = 1.0 + Dn * in2 * in2, al2 = al * al;
    r.q = al2 * al;
    r.q2 = 3.0 * al2 * be;
    r.qv = -6.0 * al2 * in2;
    r.q22 = 6.0 * al * be * be - 6.0 * al2 * Dn * in2 * in2 * in2;
    r.q2v = (-12.0 * al * be + 6.0 * al2 * in2) * in2;
    r.qvvo = 24.0 * al * in2 * in2;
    r.qvvd = small ? -6.0 * n2 : -6.0 * al2 * in2;
    return r;
}

// Phi(n) (FMTk_energy_density, src/dft_physics_FMT1.c:39-57, FMT2.c:40-58, FMT3.c:39-59, FMT4.c:38-58);
// the White Bear f's third term, 0/0 at n3 = 0, is 0 below the cut-off
template <int ND>
__host__ __device__ __attribute__((always_inline)) inline double fmt_phi(int func, const double *n)
{
    constexpr double inv_pi = 0.318309886183790671537767526745028724;
    const FmtScalars s = fmt_scalars(func, n[3]);
    double d12 = 0.0, d22 = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        d12 += n[4 + d] * n[4 + ND + d];
        d22 += n[4 + ND + d] * n[4 + ND + d];
    }
    const FmtQ q = fmt_q(func, n[2], d22);
    return n[0] * s.lnA + (n[1] * n[2] - d12) * s.a + q.q * s.f0 * (inv_pi / 36.0);
}

// dPhi/dn in n order
template <int ND>
__host__ __device__ __attribute__((always_inline)) inline void fmt_dphi(int func, const double *n, double *out)
{
    constexpr double inv_pi = 0.318309886183790671537767526745028724;
    const FmtScalars s = fmt_scalars(func, n[3]);
    double d12 = 0.0, d22 = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        d12 += n[4 + d] * n[4 + ND + d];
        d22 += n[4 + ND + d] * n[4 + ND + d];
    }
    const double a = s.a, a2 = a * a;
    const FmtQ q = fmt_q(func, n[2], d22);
    const double c0 = s.f0 * (inv_pi / 36.0), c1 = s.f1 * (inv_pi / 36.0);
    // every element is written once, selects instead of branches: the array stays in registers
    const double o2 = n[1] * a + q.q2 * c0;
    const double o3 = n[0] * a + (n[1] * n[2] - d12) * a2 + q.q * c1;
    out[0] = s.lnA;
    out[1] = n[2] * a;
    out[2] = s.cut ? 0.0 : o2;
    out[3] = s.cut ? 0.0 : o3;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        out[4 + d] = -n[4 + ND + d] * a;
        out[4 + ND + d] = s.cut ? 0.0 : (-n[4 + d] * a + q.qv * n[4 + ND + d] * c0);
    }
}

// out = H(n) dn (exact Hessian of Phi; zero below the White Bear cut-off)
template <int ND>
__host__ __device__ __attribute__((always_inline)) inline void fmt_hvp(int func, const double *n, const double *dn, double *out)
{
    constexpr double inv_pi = 0.318309886183790671537767526745028724;
    const FmtScalars s = fmt_scalars(func, n[3]);
    const bool z = s.cut;   // every element is written once, through selects (no early return)
    double d12 = 0.0, d22 = 0.0, vdv = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        d12 += n[4 + d] * n[4 + ND + d];
        d22 += n[4 + ND + d] * n[4 + ND + d];
        vdv += n[4 + ND + d] * dn[4 + ND + d];
    }
    const FmtQ q = fmt_q(func, n[2], d22);
    const double c0 = s.f0 * (inv_pi / 36.0), c1 = s.f1 * (inv_pi / 36.0), c2 = s.f2 * (inv_pi / 36.0);
    const double a = s.a, a2 = a * a, a3 = a2 * a;
    const double n0 = n[0], n1 = n[1], n2 = n[2];
    const double h22 = q.q22 * c0;
    const double h23 = n1 * a2 + q.q2 * c1;
    const double h33 = n0 * a2 + 2.0 * (n1 * n2 - d12) * a3 + q.q * c2;
    const double c2v = q.q2v * c0;   // H(2, v2d) / nv2d
    const double c3v = q.qv * c1;    // the nv2 part of H(3, v2d) / nv2d
    double o2 = a * dn[1] + h22 * dn[2] + h23 * dn[3];
    double o3 = a * dn[0] + n2 * a2 * dn[1] + h23 * dn[2] + h33 * dn[3];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        const double v1 = n[4 + d], v2 = n[4 + ND + d];
        const double dv1 = dn[4 + d], dv2 = dn[4 + ND + d];
        o2 += v2 * c2v * dv2;
        o3 += -v2 * a2 * dv1 + (-v1 * a2 + v2 * c3v) * dv2;
        const double ov1 = -v2 * a2 * dn[3] - a * dv2;
        const double ov2 = v2 * c2v * dn[2] + (-v1 * a2 + v2 * c3v) * dn[3] - a * dv1 +
                           (q.qvvd * dv2 + q.qvvo * v2 * vdv) * c0;
        out[4 + d] = z ? 0.0 : ov1;
        out[4 + ND + d] = z ? 0.0 : ov2;
    }
    out[0] = z ? 0.0 : a * dn[3];
    out[1] = z ? 0.0 : a * dn[2] + n2 * a2 * dn[3];
    out[2] = z ? 0.0 : o2;
    out[3] = z ? 0.0 : o3;
}

// The reference's delta-stencil Hessian row d2phi_drb2_delta_rb_FMTk (src/dft_physics_FMT1.c:79-126,
// FMT2.c:106-184, FMT3.c:101-160, FMT4.c:101-191) for one stencil point: out (n order) is the
// derivative of that point's Euler-Lagrange term with respect to the rho-bar unknowns at the node it
// reads, for the kernel vector vec = offset dx / R and the mirror signs sign[d] (-1 across a
// REFLECT boundary).  It is affine in vec.  It differs from the exact Hessian in the nv2 columns of
// FMT2 (2D/3D) and FMT4, and for FMT3/FMT4 across REFLECT boundaries, where the reference leaves
// the vector densities unmirrored; the reference-Jacobian mode of the product uses it to reproduce
// the reference's assembled entries.
template <int ND>
__host__ __device__ inline void fmt_ref_delta_row(int func, const double *n, const double *vec, const double *sign,
                                                  double i4pir, double i4pirsq, double *out)
{
    constexpr double kPi = 3.14159265358979323846, i4pi = 1.0 / (4.0 * kPi);
    const double a = 1.0 / (1.0 - n[3]), a2 = a * a, a3 = a2 * a;
    if (func == 0 || func == 1) {   // Rosenfeld and FMT2: mirrored vector densities
        double n1v[ND > 0 ? ND : 1], n2v[ND > 0 ? ND : 1];
        double d22 = 0.0;
        for (int d = 0; d < ND; ++d) {
            n2v[d] = sign[d] * n[4 + ND + d];
            n1v[d] = sign[d] * n[4 + d];
            d22 += n2v[d] * n2v[d];
        }
        if (func == 0) {
            double S2 = a * i4pir + n[2] * i4pi * a2, S3 = i4pirsq * a + n[2] * i4pir * a2 + n[1] * a2 + n[2] * n[2] * i4pi * a3;
            for (int d = 0; d < ND; ++d) {
                S2 += vec[d] * n2v[d] * i4pi * a2;
                S3 += a2 * (-n2v[d] * n2v[d] * i4pi * a + vec[d] * (n2v[d] * i4pir + n1v[d] + n[2] * n2v[d] * a / (2.0 * kPi)));
            }
            out[0] = 0.0;
            out[1] = a;
            out[2] = S2;
            out[3] = S3;
            for (int d = 0; d < ND; ++d) {
                out[4 + d] = sign[d] * a * vec[d];
                out[4 + ND + d] = sign[d] * (-n2v[d] * i4pi * a2 + (i4pir * a + n[2] * i4pi * a2) * vec[d]);
            }
            return;
        }
        double al, be, ep, gam[ND > 0 ? ND : 1], dg = 0.0;
        if (n[2] > 1e-15) {
            al = n[2] - d22 / n[2];
            be = 1.0 + d22 / (n[2] * n[2]);
            for (int d = 0; d < ND; ++d) {
                gam[d] = n2v[d] / n[2];
                dg += gam[d] * gam[d];
            }
            ep = al / n[2];
        } else {
            al = n[2];
            be = 1.0;
            for (int d = 0; d < ND; ++d) gam[d] = 0.0;
            ep = 1.0;
        }
        double S2 = a * i4pir + i4pi * a2 * al * (be * be - ep * dg);
        double S3 = i4pirsq * a + n[2] * i4pir * a2 + n[1] * a2 + i4pi * a3 * al * al * be;
        for (int d = 0; d < ND; ++d) {
            S2 -= i4pi * a2 * al * (ep - 2.0 * be) * gam[d] * vec[d];
            S3 += a2 * vec[d] * (n2v[d] * i4pir + n1v[d]) + 2.0 * i4pi * a3 * al * al * gam[d] * vec[d];
        }
        out[0] = 0.0;
        out[1] = a;
        out[2] = S2;
        out[3] = S3;
        for (int d = 0; d < ND; ++d) {
            out[4 + d] = sign[d] * a * vec[d];
            out[4 + ND + d] = sign[d] * (i4pir * a * vec[d] + i4pi * a2 * al * (ep - 2.0 * be) * gam[d] -
                                         i4pi * a2 * al * (4.0 * dg - ep) * vec[d]);
        }
        return;
    }
    // White Bear and FMT4: the vector densities enter unmirrored, the signs only on the V columns
    for (int q = 0; q < 4 + 2 * ND; ++q) out[q] = 0.0;
    if (!(n[3] > 1e-10)) return;
    double d22 = 0.0, dnw = 0.0;
    for (int d = 0; d < ND; ++d) {
        d22 += n[4 + ND + d] * n[4 + ND + d];
        dnw += n[4 + ND + d] * vec[d];
    }
    const double n3sq = n[3] * n[3], n3cb = n3sq * n[3], l = log(1.0 - n[3]);
    const double fac1 = n[3] - 2.0 * (1.0 - n[3]) * l, fac2 = n[3] + (1.0 - n[3]) * (1.0 - n[3]) * l;
    const double fac3 = -(a2 / n3cb) + (a3 / n3sq), g = 2.0 * fac2 * fac3 + fac1 * a2 / n3sq;
    const double h = fac2 * a2 / (6.0 * kPi * n3sq);
    if (func == 2) {
        double S2 = a * i4pir + n[2] * h, S3 = i4pirsq * a + i4pir * n[2] * a2 + n[1] * a2 + (n[2] * n[2] - d22) * g / (12.0 * kPi);
        for (int d = 0; d < ND; ++d) {
            S2 += vec[d] * n[4 + ND + d] * h;
            S3 += vec[d] * (n[4 + ND + d] * i4pir * a2 + n[4 + d] * a2 + n[2] * n[4 + ND + d] * g / (6.0 * kPi));
        }
        out[1] = a;
        out[2] = S2;
        out[3] = S3;
        for (int d = 0; d < ND; ++d) {
            out[4 + d] = sign[d] * a * vec[d];
            out[4 + ND + d] = sign[d] * (-n[4 + ND + d] * h + (i4pir * a + n[2] * h) * vec[d]);
        }
        return;
    }
    const double fac = n[2] > 1e-15 ? 1.0 - d22 / (n[2] * n[2]) : 1.0, i2 = n[2] > 1e-15 ? 1.0 / n[2] : 0.0;
    double S2 = a * i4pir + h * (n[2] * fac * fac * fac + 3.0 * d22 * fac * fac * i2 + 4.0 * d22 * d22 * fac * i2 * i2 * i2);
    double S3 = i4pirsq * a + i4pir * n[2] * a2 + n[1] * a2 + (n[2] * n[2] * fac * fac * fac + 2.0 * d22 * fac * fac) * g / (12.0 * kPi);
    for (int d = 0; d < ND; ++d) {
        S2 += vec[d] * n[4 + ND + d] * h * (fac * fac + 4.0 * d22 * fac * i2 * i2);
        S3 += vec[d] * (n[4 + ND + d] * i4pir * a2 + n[4 + d] * a2 + n[2] * n[4 + ND + d] * fac * fac * g / (6.0 * kPi));
    }
    out[1] = a;
    out[2] = S2;
    out[3] = S3;
    for (int d = 0; d < ND; ++d) {
        out[4 + d] = sign[d] * a * vec[d];
        out[4 + ND + d] = sign[d] * (-n[4 + ND + d] * h * (fac * fac - 4.0 * d22 * fac * i2 + 4.0 * dnw * fac * i2) +
                                     (i4pir * a + n[2] * fac * fac * h) * vec[d]);
    }
}

// Adjoint fields of the Jacobian-vector product in the reference-Jacobian mode for component c:
// A = t0 . dn and B_d = -(t_d . dn), where the reference's delta row (fmt_ref_delta_row) is
// t0 + sum_d vec_d t_d; the Euler-Lagrange sum then reproduces the reference's assembled entries
// (sum_k w_k row(vec_k) . dn at the node that stencil point k reads)
template <int ND>
__host__ __device__ inline void fmt_ref_adjoint(int func, const double *n, const double *dn, const double *sign,
                                                double i4pir, double i4pirsq, double &A, double *B)
{
    constexpr int NR = 4 + 2 * ND;
    double vec[ND > 0 ? ND : 1], row[NR];
    for (int d = 0; d < ND; ++d) vec[d] = 0.0;
    fmt_ref_delta_row<ND>(func, n, vec, sign, i4pir, i4pirsq, row);
    A = 0.0;
    for (int q = 0; q < NR; ++q) A += row[q] * dn[q];
    for (int d = 0; d < ND; ++d) {
        vec[d] = 1.0;
        double r1[NR];
        fmt_ref_delta_row<ND>(func, n, vec, sign, i4pir, i4pirsq, r1);
        vec[d] = 0.0;
        double t = 0.0;
        for (int q = 0; q < NR; ++q) t += (r1[q] - row[q]) * dn[q];
        B[d] = -t;
    }
}

}  // namespace tmi
