// swe_kernels.h -- device-side arithmetic of the explicit shallow-water step, shared by every
// kernel.  Each function restates one reference routine with the reference's evaluation order
// (the file is compiled with -ffp-contract=off, so no multiply-add is fused):
//   ovh_of         compute_one_over_h          SOURCES/condlim.f90:79-90
//   phi_f          phi                         SOURCES/condlim.f90:960-981
//   lambda_vacc    compute_lambda_vacc         SOURCES/condlim.f90:879-957
//   mu_of          compute_muij (one entry)    SOURCES/update.f90:338-356
//   pow43          MAX(h,htiny)**(4/3) of friction, SOURCES/update.f90:437
// Fortran MAX(a,b) is (a > b ? a : b) and MIN(a,b) is (a < b ? a : b), as flang lowers them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

// One node of a state buffer: (h, qx, qy) and the bathymetry z, 32 bytes, so a neighbour gather
// is one aligned 32-byte load.
struct alignas(32) Node {
    double h, qx, qy, z;
};

// Scalars fixed per parameter set, computed on the host with the same expressions the
// reference evaluates at run time (host sqrt is correctly rounded).
struct Consts {
    double g, htiny, cfl, max_water_h, mannings;
    double x0;          // (2*SQRT(2)-1)**2
    double a_c2;        // 1/(2*SQRT(2))
    double two_sqrt_g;  // 2*SQRT(g)
    double sqrt_2_g;    // SQRT(2/g)
    double g_half;      // g/2
    double g16;         // 16*g
    double fric;        // g*mannings**2
    double ent_cst;     // 1e-10*g*max_water_h**2
    int type_test;
    int check_h;        // check_hmin is active for type_test 1..13
};

enum : int { FLAG_NEG_H = 1, FLAG_LAMBDA = 2 };

__device__ __forceinline__ double fmaxF(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double fminF(double a, double b) { return (a < b) ? a : b; }

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
    const double Sl = sqrt(gl ? K.g_half * (ovh + ovhl) : gh);
    const double Sr = sqrt(gr ? K.g_half * (ovh + ovhr) : gh);
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
        const double sqrmin = sqrt(hmin);
        const double sqrmax = sqrt(hmax);
        if (0.0 <= fh2) {
            const double cc = -hmin * K.a_c2 - sqrmin * sqrmax + sqrmin * (vr - vl) / K.two_sqrt_g;
            const double Delta = hmin - 4.0 * K.a_c2 * cc;
            if (Delta < 0.0) {
                err |= FLAG_LAMBDA;
                return 0.0;
            }
            const double t = (-sqrmin + sqrt(Delta)) / (2.0 * K.a_c2);
            ht = fminF(K.x0 * hmax, t * t);
        } else {
            ht = sqrmin * sqrmax * (1.0 + K.sqrt_2_g * (vl - vr) / (sqrmin + sqrmax));
        }
    }
    const double lbdl = vl - sql * sqrt((1.0 + fmaxF((ht - hl) * ovhl / 2.0, 0.0)) *
                                        (1.0 + fmaxF((ht - hl) * ovhl, 0.0)));
    const double lbdr = vr + sqr * sqrt((1.0 + fmaxF((ht - hr) * ovhr / 2.0, 0.0)) *
                                        (1.0 + fmaxF((ht - hr) * ovhr, 0.0)));
    return fmaxF(fabs(lbdl), fabs(lbdr));
}

// First-order graph viscosity of one off-diagonal entry: |c_ij| * lambda(U_i, U_j, c_ij/|c_ij|)
// (compute_dij, SOURCES/update.f90:305-315).  norm = SQRT(SUM(c_ij**2)) and n = c_ij/norm depend
// only on the mesh and are evaluated once on the host with the same expressions.
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

// XCD-aware block order: MI355X dispatches workgroup b to XCD (b mod 8); each XCD gets a
// contiguous range of rows, so neighbouring rows share that XCD's L2.
__device__ __forceinline__ int xcd_block() {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, w = b >> 3;
    return x * q + (x < r ? x : r) + w;
}

}  // namespace swe
