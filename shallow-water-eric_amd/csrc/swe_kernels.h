// swe_kernels.h -- device-side arithmetic of the explicit shallow-water step, shared by every
// kernel.  Each function restates one reference routine with the reference's evaluation order
// (the file is compiled with -ffp-contract=off, so no multiply-add is fused):
//   ovh_of         compute_one_over_h          SOURCES/condlim.f90:79-90
//   phi_f          phi                         SOURCES/condlim.f90:960-981
//   lambda_vacc    compute_lambda_vacc         SOURCES/condlim.f90:879-957
//   mu_of          compute_muij (one entry)    SOURCES/update.f90:338-356
//   pow43          MAX(h,htiny)**(4/3) of friction, SOURCES/update.f90:437
//   sin_rn         sin rounded to nearest, for the inflow boundary value
// Fortran MAX(a,b) is (a > b ? a : b) and MIN(a,b) is (a < b ? a : b), as flang lowers them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swe_layout.h"

namespace swe {

__device__ __forceinline__ double fmaxF(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double fminF(double a, double b) { return (a < b) ? a : b; }

// SQRT of an argument that is a finite double of at least 2^-767: the instructions of the correctly
// rounded sqrt hipcc emits for gfx950 (reciprocal square root, then Newton corrections in FMA),
// without its scaling of tiny arguments and its handling of 0 and infinity, which such arguments
// do not need -- the same bits for every such argument (checked exhaustively on random arguments).
// Used where the argument is >= 1 (lambda's wave speeds) or a squared |c_ij| of an off-diagonal
// entry (zero only on a degenerate mesh, on which the reference's d_ij is NaN).
__device__ __forceinline__ double sqrt_normal(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return g;
}

// SQRT, the same bits as sqrt() for every argument: sqrt_normal, and sqrt() itself for the lanes
// whose argument is not a finite double >= 2^-767 (zero on dry nodes, negative, infinite, NaN), a
// branch only the waves that have such a lane take.
__device__ __forceinline__ double sqrt_rn(double x) {
    double r = sqrt_normal(x);
    if (__builtin_expect(!(x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023), 0)) r = sqrt(x);
    return r;
}

__device__ __forceinline__ double ovh_of(double h, double htiny) {
    const double m = fmaxF(h, htiny);
    return 2.0 * h / (h * h + m * m);
}

// phi(h, ul, hl, ur, hr); ovhl, ovhr are compute_one_over_h of hl, hr and sql, sqr are
// SQRT(g*ABS(hl)), SQRT(g*ABS(hr)).  Each side takes one square root of the argument its branch
// uses, so a wave never runs both branches: fl = (h-hl)*SQRT((g/2)*(ovh+ovhl)) if h > hl, else
// 2*(SQRT(g*h) - SQRT(g*hl)), where SQRT(g*hl) is sql for hl >= 0 (NaN otherwise, as in the reference).
__device__ __forceinline__ double phi_f(double h, double ul, double hl, double ur, double hr,
                                        double ovhl, double ovhr, double sql, double sqr, const Consts& K) {
    const bool gl = h > hl, gr = h > hr;
    const double ovh = ovh_of(h, K.htiny);
    const double gh = K.g * h;
    const double Sl = sqrt_rn(gl ? K.g_half * (ovh + ovhl) : gh);
    const double Sr = sqrt_rn(gr ? K.g_half * (ovh + ovhr) : gh);
    const double nan = __builtin_nan("");
    const double sgl = (hl >= 0.0) ? sql : nan;
    const double sgr = (hr >= 0.0) ? sqr : nan;
    const double fl = gl ? (h - hl) * Sl : 2.0 * (Sl - sgl);
    const double fr = gr ? (h - hr) * Sr : 2.0 * (Sr - sgr);
    return fl + fr + ur - ul;
}

// compute_lambda_vacc for the left state (hl, qxl, qyl) and the right state (hr, qxr, qyr) along
// the unit vector (n1, n2); ovhl/ovhr and sql = SQRT(g*ABS(hl)), sqr = SQRT(g*ABS(hr)) are passed
// in because the callers already hold them.
__device__ __forceinline__ double lambda_vacc(double hl, double qxl, double qyl, double hr,
                                              double qxr, double qyr, double n1, double n2,
                                              double ovhl, double ovhr, double sql, double sqr,
                                              const Consts& K, int& err) {
    const double vl = (0.0 + qxl * n1 + qyl * n2) * ovhl;
    const double vr = (0.0 + qxr * n1 + qyr * n2) * ovhr;
    double hmin, hmax;
    if (hl <= hr) { hmin = hl; hmax = hr; }
    else { hmin = hr; hmax = hl; }
    double ht;
    const double fh = phi_f(K.x0 * hmin, vl, hl, vr, hr, ovhl, ovhr, sql, sqr, K);
    if (0.0 <= fh) {
        const double t = fmaxF(vl - vr + 2.0 * sql + 2.0 * sqr, 0.0);
        ht = fminF(K.x0 * hmin, t * t / K.g16);
    } else {
        const double fh2 = phi_f(K.x0 * hmax, vl, hl, vr, hr, ovhl, ovhr, sql, sqr, K);
        const double sqrmin = sqrt_rn(hmin);
        const double sqrmax = sqrt_rn(hmax);
        if (0.0 <= fh2) {
            const double cc = -hmin * K.a_c2 - sqrmin * sqrmax + sqrmin * (vr - vl) / K.two_sqrt_g;
            const double Delta = hmin - 4.0 * K.a_c2 * cc;
            if (Delta < 0.0) {
                err |= FLAG_LAMBDA;
                return 0.0;
            }
            const double t = (-sqrmin + sqrt_rn(Delta)) / (2.0 * K.a_c2);
            ht = fminF(K.x0 * hmax, t * t);
        } else {
            ht = sqrmin * sqrmax * (1.0 + K.sqrt_2_g * (vl - vr) / (sqrmin + sqrmax));
        }
    }
    // the arguments are products of two factors >= 1
    const double lbdl = vl - sql * sqrt_normal((1.0 + fmaxF((ht - hl) * ovhl / 2.0, 0.0)) *
                                               (1.0 + fmaxF((ht - hl) * ovhl, 0.0)));
    const double lbdr = vr + sqr * sqrt_normal((1.0 + fmaxF((ht - hr) * ovhr / 2.0, 0.0)) *
                                               (1.0 + fmaxF((ht - hr) * ovhr, 0.0)));
    return fmaxF(fabs(lbdl), fabs(lbdr));
}

// First-order graph viscosity of one off-diagonal entry: |c_ij| * lambda(U_i, U_j, c_ij/|c_ij|)
// (compute_dij, SOURCES/update.f90:305-315), given norm = SQRT(SUM(c_ij**2)) and n = c_ij/norm.
__device__ __forceinline__ double dij_of(double2 n, double norm, const Node& I, const Node& J,
                                         double ovh_i, double ovh_j, double sql, double sqr,
                                         const Consts& K, int& err) {
    const double lam = lambda_vacc(I.h, I.qx, I.qy, J.h, J.qx, J.qy, n.x, n.y, ovh_i, ovh_j, sql, sqr, K, err);
    return norm * lam;
}

// compute_muij of one off-diagonal entry; (v1i, v2i) and (v1j, v2j) are q*one_over_h.
__device__ __forceinline__ double mu_of(double cx, double cy, double v1i, double v2i, double v1j, double v2j) {
    const double sl = 0.0 + cx * v1i + cy * v2i;
    const double sr = 0.0 + cx * v1j + cy * v2j;
    return fmaxF(fmaxF(-sl, 0.0), fmaxF(sr, 0.0));
}

// ---- x**(4/3) rounded to nearest, for the Manning friction term -------------------------------
// The reference evaluates MAX(h,htiny)**(4.d0/3.d0) with libm pow, which is correctly rounded
// except in rare near-halfway cases.  Here x^(4/3) = x*cbrt(x) is evaluated in double-double
// arithmetic (Dekker products via fma, Newton steps for the cube root) and rounded once, which
// gives the correctly rounded result apart from exact ties.
struct dd_t { double hi, lo; };

__host__ __device__ __forceinline__ dd_t dd_two_prod(double a, double b) {
    const double p = a * b;
    return {p, fma(a, b, -p)};
}
__host__ __device__ __forceinline__ dd_t dd_add(dd_t a, dd_t b) {
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double e = (a.hi - (s - bb)) + (b.hi - bb);
    const double lo = e + a.lo + b.lo;
    const double hi = s + lo;
    return {hi, lo - (hi - s)};
}
__host__ __device__ __forceinline__ dd_t dd_mul(dd_t a, dd_t b) {
    dd_t p = dd_two_prod(a.hi, b.hi);
    const double lo = p.lo + (a.hi * b.lo + a.lo * b.hi);
    const double hi = p.hi + lo;
    return {hi, lo - (hi - p.hi)};
}
__host__ __device__ __forceinline__ dd_t dd_div(dd_t a, dd_t b) {
    const double q1 = a.hi / b.hi;
    dd_t r = dd_add(a, dd_mul({-q1, 0.0}, b));
    const double q2 = r.hi / b.hi;
    const double hi = q1 + q2;
    return {hi, q2 - (hi - q1)};
}

__host__ __device__ __forceinline__ double pow43(double x) {
    if (!(x > 0.0) || isinf(x)) return pow(x, 4.0 / 3.0);
    // scale to [1, 8): x = m * 2^(3k)
    int e;
    double m = frexp(x, &e);          // x = m*2^e, m in [0.5,1)
    int k = (e - 1);
    int r = ((k % 3) + 3) % 3;
    k = (k - r) / 3;                  // x = (m*2^(r+1)) * 2^(3k)
    m = ldexp(m, r + 1);              // m in [1, 8)
    // cube root of m in double-double: y <- y - (y^3 - m)/(3 y^2), twice in dd
    double y0 = cbrt(m);
    dd_t y = {y0, 0.0};
    dd_t M = {m, 0.0};
    for (int it = 0; it < 2; ++it) {
        dd_t y2 = dd_mul(y, y);
        dd_t y3 = dd_mul(y2, y);
        dd_t num = dd_add(y3, {-M.hi, -M.lo});
        dd_t den = dd_mul({3.0, 0.0}, y2);
        dd_t corr = dd_div(num, den);
        y = dd_add(y, {-corr.hi, -corr.lo});
    }
    dd_t p = dd_mul(M, y);            // m^(4/3)
    // x^(4/3) = m^(4/3) * 2^(4k); the final rounding happens in hi+lo below
    return ldexp(p.hi + p.lo, 4 * k);
}

// ---- sin(x) rounded to nearest, for the inflow boundary value --------------------------------
// The host's libm sin (called where the boundary value is evaluated on the CPU) returns the
// correctly rounded result in all but rare near-halfway cases.  sin_rn evaluates the same value
// on the device: k = nearest integer to x*2/pi, r = x - k*pi/2 in double-double with pi/2 split
// into three doubles (k*P0 is exact through fma, and x - k*P0 is exact by Sterbenz' lemma), then
// the Taylor series of sin(r) or cos(r) (|r| <= pi/4, 15 terms, coefficients (-1)^n/(2n+1)! and
// (-1)^n/(2n)! as double-double) in double-double Horner form, rounded once.  The error before
// the final rounding is below 2^-100 relative for |x| < 2^29.
__host__ __device__ inline double sin_rn(double x) {
    if (!(fabs(x) < 536870912.0)) return sin(x);     // outside the reduction's range; NaN, inf
    const double k = rint(x * 0.6366197723675814);   // 2/pi
    const dd_t p0 = dd_two_prod(k, 1.5707963267948966);
    const dd_t p1 = dd_two_prod(k, 6.123233995736766e-17);
    dd_t r = {x - p0.hi, 0.0};
    r = dd_add(r, {-p0.lo, 0.0});
    r = dd_add(r, {-p1.hi, -p1.lo});
    r = dd_add(r, {-(k * -1.4973849048591698e-33), 0.0});
    const dd_t r2 = dd_mul(r, r);
    const int q = (int)((long long)k & 3);
    double v;
    if (q & 1) {
        dd_t s = {3.279889237069838e-30, 1.5117542744029879e-46};
        s = dd_add(dd_mul(s, r2), {-2.4795962632247976e-27, 1.2953730964765229e-43});
        s = dd_add(dd_mul(s, r2), {1.6117375710961184e-24, -3.6846573564509766e-41});
        s = dd_add(dd_mul(s, r2), {-8.896791392450574e-22, 7.911402614872376e-38});
        s = dd_add(dd_mul(s, r2), {4.110317623312165e-19, 1.4412973378659527e-36});
        s = dd_add(dd_mul(s, r2), {-1.5619206968586225e-16, -1.1910679660273754e-32});
        s = dd_add(dd_mul(s, r2), {4.779477332387385e-14, 4.399205485834081e-31});
        s = dd_add(dd_mul(s, r2), {-1.1470745597729725e-11, -2.0655512752830745e-28});
        s = dd_add(dd_mul(s, r2), {2.08767569878681e-09, -1.20734505911326e-25});
        s = dd_add(dd_mul(s, r2), {-2.755731922398589e-07, -2.3767714622250297e-23});
        s = dd_add(dd_mul(s, r2), {2.48015873015873e-05, 2.1511947866775882e-23});
        s = dd_add(dd_mul(s, r2), {-0.001388888888888889, 5.300543954373577e-20});
        s = dd_add(dd_mul(s, r2), {0.041666666666666664, 2.3129646346357427e-18});
        s = dd_add(dd_mul(s, r2), {-0.5, 0.0});
        s = dd_add(dd_mul(s, r2), {1.0, 0.0});
        v = s.hi + s.lo;
    } else {
        dd_t s = {1.1309962886447716e-31, 1.0498015412959506e-47};
        s = dd_add(dd_mul(s, r2), {-9.183689863795546e-29, -1.4303150396787322e-45});
        s = dd_add(dd_mul(s, r2), {6.446950284384474e-26, -1.9330404233703465e-42});
        s = dd_add(dd_mul(s, r2), {-3.868170170630684e-23, 8.843177655482344e-40});
        s = dd_add(dd_mul(s, r2), {1.9572941063391263e-20, -1.3643503830087908e-36});
        s = dd_add(dd_mul(s, r2), {-8.22063524662433e-18, -2.2141894119604265e-34});
        s = dd_add(dd_mul(s, r2), {2.8114572543455206e-15, 1.6508842730861433e-31});
        s = dd_add(dd_mul(s, r2), {-7.647163731819816e-13, -7.03872877733453e-30});
        s = dd_add(dd_mul(s, r2), {1.6059043836821613e-10, 1.2585294588752098e-26});
        s = dd_add(dd_mul(s, r2), {-2.505210838544172e-08, 1.448814070935912e-24});
        s = dd_add(dd_mul(s, r2), {2.7557319223985893e-06, -1.858393274046472e-22});
        s = dd_add(dd_mul(s, r2), {-0.0001984126984126984, -1.7209558293420705e-22});
        s = dd_add(dd_mul(s, r2), {0.008333333333333333, 1.1564823173178714e-19});
        s = dd_add(dd_mul(s, r2), {-0.16666666666666666, -9.25185853854297e-18});
        s = dd_add(dd_mul(s, r2), {1.0, 0.0});
        s = dd_mul(s, r);
        v = s.hi + s.lo;
    }
    return (q & 2) ? -v : v;
}

// XCD-aware block order: MI355X dispatches workgroup b to XCD (b mod 8); each XCD gets a
// contiguous range of rows, so neighbouring rows share that XCD's L2.
__device__ __forceinline__ int xcd_block() {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, w = b >> 3;
    return x * q + (x < r ? x : r) + w;
}

}  // namespace swe
