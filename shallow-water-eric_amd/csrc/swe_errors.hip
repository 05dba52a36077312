// swe_errors.hip -- compute_errors (SOURCES/main.f90:198-263) on the device state: sol_anal
// (SOURCES/condlim.f90:311-831, analytic tests 3 to 7, 11, 12 and 13) at the Gauss points of r_gauss
// (main.f90:303-322), and the quadrature sums of ns_anal_l1 / ns_anal_0
// (SFEMaNS/FEMSUB_OBJET/fem_tn.f90:1679-1716, 1513-1553) of u(k,:) - exact and 0 - exact; and the
// time loop's unit-10 diagnostic of tests 1, 2, 10 (main.f90:94-99), ns_l1 (fem_tn.f90:2119-2153) of qx,
// qy and h*sqrt(h) with the ratio 0.5*(q1+q2)/(g*q3).
//
// Every summand is computed with the reference's expressions (the file is compiled without FMA
// contraction; the factors of sol_anal that depend on t only are evaluated once on the host with the
// C library's cos, sin, exp, acos and pow, as the reference's runtime does).  The reference sums the summands
// sequentially in element order; here each sum is accumulated in double-double arithmetic (error-free
// two-sums) over a fixed partition of the elements and reduced in a fixed order, which gives the
// exactly rounded sum up to a relative 2^-100 or so, the same on every run.  It differs from the
// sequential sum by that sum's own rounding error, at most (n-1)u * (sum of |summands|).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "swe_engine.h"
#include "swe_p1.h"

namespace swe {

namespace {

// t-dependent factors of sol_anal, evaluated in the reference's order (condlim.f90 lines cited).
struct Anal {
    int test = 0;
    double htiny = 0.0, t = 0.0;
    // test 4 (385-428)
    int small_t = 0;
    double x0 = 0.0, hL = 0.0, speed = 0.0, neg_ts = 0.0, two_ts = 0.0, two_t = 0.0, c49 = 0.0, c827 = 0.0;
    // tests 5, 11, 12: per-point expressions are X - xc, bathi's h0 and a*a, and the factors below
    double xc = 0.0, h0 = 0.0, a2 = 0.0;
    double k05 = 0.0, cs = 0.0, sn = 0.0, cm = 0.0, cp = 0.0;                   // test 5 (429-456)
    double A1 = 0.0, cB = 0.0, P1 = 0.0, P2 = 0.0, gam = 0.0, E2 = 0.0;           // test 11 (649-700)
    double cT1 = 0.0, cT2 = 0.0, T3 = 0.0, delta = 0.0, eta = 0.0, w = 0.0;       // test 12 (701-750)
    // the steady flows over the bump, tests 3 (346-384), 6 (457-521) and 7 (522-605)
    double q0 = 0.0, b0 = 0.0, d = 0.0, ber = 0.0, fpio3 = 0.0, xshock = 0.0, bath_shock = 0.0, h_post = 0.0;
    // test 13, the SGN model's solitary wave (751-827): h1, h2, crest x0, celerity D, z and D*t
    double h1 = 0.0, h2 = 0.0, zz = 0.0, Dw = 0.0, Dt = 0.0;
};

// The bump of tests 2, 3, 6, 7: MAX(0, 0.2*(x-8)**3*(12-x)**3/64) (condlim.f90:355).
__host__ __device__ inline double bump_of(double x) {
    const double a = x - 8.0, c = 12.0 - x;
    const double v = 0.2 * (a * a * a) * (c * c * c) / 64.0;
    return (0.0 > v) ? 0.0 : v;
}
inline double bump_host(double x) { return bump_of(x); }

int make_anal(int test, double t, double g, double htiny, Anal& A) {
    A = Anal{};
    A.test = test;
    A.htiny = htiny;
    A.t = t;
    if (test == 4) {
        A.x0 = 5.0;
        A.hL = 0.005;
        A.speed = std::sqrt(g * A.hL);
        A.small_t = (t <= 1.0e-14) ? 1 : 0;
        A.neg_ts = -(t * A.speed);
        A.two_ts = 2.0 * t * A.speed;
        A.two_t = 2.0 * t;
        A.c49 = 4.0 / (9.0 * g);
        A.c827 = 8.0 / (27.0 * g);
        return SWE_OK;
    }
    if (test == 3) {
        A.small_t = (t <= 1.0e-14) ? 1 : 0;
        A.q0 = 4.42;
        A.hL = 2.0;
        A.b0 = (A.q0 * A.q0) / (2.0 * g * (A.hL * A.hL));
        A.d = (A.q0 * A.q0) / (2.0 * g);
        return SWE_OK;
    }
    if (test == 6 || test == 7) {
        A.small_t = (t <= 1.0e-14) ? 1 : 0;
        A.x0 = 10.0;
        A.hL = 0.66;
        A.q0 = (test == 6) ? 1.53 : 0.18;
        const double tpio3 = 2.0 * std::acos(-1.0) / 3.0;
        A.fpio3 = 2.0 * tpio3;
        A.d = (A.q0 * A.q0) / (2.0 * g);
        const double eta = -std::pow(27.0 * A.d / 4.0, 1.0 / 3.0);
        A.ber = 0.2 - eta;                                        // bernoulli (test 6), Ber_Pre (test 7)
        if (test == 7) {
            A.xshock = 11.7;
            if (8.0 < A.xshock && A.xshock < 12.0) A.bath_shock = bump_host(A.xshock);
            const double b = A.bath_shock - A.ber;
            const double R = (-(27.0 * A.d) - 2.0 * (b * b * b)) / 54.0, Q = -((b * b) / 9.0);
            if (!(Q * Q * Q + R * R >= 0.0)) {
                const double theta = std::acos(R / std::sqrt(-(Q * Q * Q)));
                const double x3 = 2.0 * std::sqrt(-Q) * std::cos(theta / 3.0 + A.fpio3) - b / 3.0;
                A.h_post = (-x3 + std::sqrt(x3 * x3 + 8.0 * (A.q0 * A.q0) / (g * x3))) / 2.0;
            }
        }
        return SWE_OK;
    }
    if (test == 5) {
        const double a = 1.0, eta = 0.5;
        A.h0 = 0.1;
        A.xc = 4.0 / 2.0;
        A.a2 = a * a;
        const double omega = std::sqrt(2.0 * g * A.h0) / a;
        A.k05 = eta * A.h0 / A.a2;
        A.cs = std::cos(omega * t);
        A.sn = std::sin(omega * t);
        A.cm = eta * omega * A.sn;   // -eta*omega*SIN(omega*t)*vv = -((eta*omega*S)*vv)
        A.cp = eta * omega * A.cs;
        return SWE_OK;
    }
    if (test == 11) {   // the case sets gravity to 9.81
        const double gg = 9.81, a = 1.0, kappa = 0.2, gamma = 0.5;
        A.h0 = 0.1;
        A.xc = 4.0 / 2.0;
        A.a2 = a * a;
        const double p = std::sqrt((8.0 * gg * A.h0) / A.a2);
        const double s = 1.0 / 2.0 * std::sqrt(p * p - kappa * kappa);
        const double c = std::cos(s * t), sn = std::sin(s * t);
        A.A1 = (gamma * gamma) / (2.0 * gg) * std::exp(-(kappa * t));
        A.E2 = std::exp(-(kappa * t / 2.0));
        A.cB = gamma / (2.0 * gg) * A.E2;
        A.P1 = 2.0 * s * c + kappa * sn;
        A.P2 = kappa * c - 2.0 * s * sn;
        A.gam = gamma;
        A.cs = c;
        A.sn = sn;
        return SWE_OK;
    }
    if (test == 12) {   // the case sets gravity to 9.81
        const double gg = 9.81, a = 3.0, f = 2.0, eta = 0.5;
        A.h0 = 0.5;
        A.xc = 0.0 / 2.0;
        const double a2 = a * a, a4 = a2 * a2, a6 = a4 * a2, a8 = a4 * a4, f2 = f * f, f4 = f2 * f2;
        A.a2 = a2;
        const double den = f2 * a4 + std::sqrt(8.0 * f2 * gg * A.h0 * a6 + f4 * a8);
        A.delta = 2.0 * f * a4 / den;
        A.w = -(1.0 / (std::sqrt(2.0) * a2) * std::sqrt(4.0 * gg * A.h0 * a2 + den));
        A.cs = std::cos(A.w * t);
        A.sn = std::sin(A.w * t);
        A.cT1 = eta / gg * A.cs * (1.0 - A.delta * f);
        A.cT2 = eta / gg * A.sn * (f / A.w - A.delta * A.w);
        A.T3 = eta * eta / (2.0 * gg) * (A.delta * A.delta * A.cs * A.cs + A.sn * A.sn / (A.w * A.w));
        A.eta = eta;
        return SWE_OK;
    }
    if (test == 13) {   // the case sets gravity to 9.81; h, qx, qy are defined at every t
        const double gg = 9.81;
        A.h1 = 10.0 / 100.0;
        A.h2 = 11.0 / 100.0;
        A.x0 = 2.75;
        A.Dw = std::sqrt(gg * A.h2);
        A.zz = std::sqrt((3.0 * (A.h2 - A.h1)) / (A.h2 * (A.h1 * A.h1)));
        A.Dt = A.Dw * t;
        return SWE_OK;
    }
    return fail(SWE_EBADOPT,
                "compute_errors: no analytic solution for type_test %d (tests 3 to 7, 11, 12 and 13 have one)", test);
}

__device__ __forceinline__ double fmaxF(double a, double b) { return (a > b) ? a : b; }

// sol_anal(1..3) at one point, given the t-dependent factors.  The steady bump flows take the roots
// of a cubic per point with the device's acos, cos and pow, and the SGN wave the device's cosh, which
// may differ from the host C library's in the last bit.
__device__ __forceinline__ void sol_anal3(const Anal& A, double x, double y, double v[3]) {
    if (A.test == 13) {
        // h1 + (h2-h1)*(1/COSH(z/2*(x-x0-D*t)))**2, h*D - h1*D, h*0 (condlim.f90:759-798)
        const double sech = 1.0 / cosh(1.0 / 2.0 * A.zz * (x - A.x0 - A.Dt));
        const double vv = fmaxF(A.h1 + (A.h2 - A.h1) * (sech * sech), 0.0);
        v[0] = vv;
        v[1] = vv * A.Dw - A.h1 * A.Dw;
        v[2] = vv * 0.0;
        return;
    }
    if (A.test == 3 || A.test == 6 || A.test == 7) {
        v[2] = 0.0;
        const double bathi = bump_of(x);
        if (A.small_t) {
            v[0] = ((A.test == 7) ? A.h_post : A.hL) - bathi;
            v[1] = (A.test == 7) ? A.q0 : 0.0;
            return;
        }
        v[1] = A.q0;
        const double b = (A.test == 3) ? bathi - A.b0 - A.hL : bathi - A.ber;
        const double R = (-(27.0 * A.d) - 2.0 * (b * b * b)) / 54.0, Q = -((b * b) / 9.0);
        const double D = Q * Q * Q + R * R;
        if (A.test != 3 && D >= 0.0) {
            const double S = -pow(-sqrt(D) - R, 1.0 / 3.0);
            const double T = -pow(sqrt(D) - R, 1.0 / 3.0);
            v[0] = -((S + T) / 2.0) - b / 3.0;
            return;
        }
        const double theta = acos(R / sqrt(-(Q * Q * Q)));
        const double x1 = 2.0 * sqrt(-Q) * cos(theta / 3.0) - b / 3.0;
        if (A.test == 3) {
            v[0] = x1;
            return;
        }
        const double x3 = 2.0 * sqrt(-Q) * cos(theta / 3.0 + A.fpio3) - b / 3.0;
        if (x < A.x0) v[0] = (x1 > x3) ? x1 : x3;
        else if (A.test == 6 || x < A.xshock) v[0] = (x1 < x3) ? x1 : x3;
        else v[0] = A.bath_shock + A.h_post - bathi;
        return;
    }
    if (A.test == 4) {
        v[2] = 0.0;
        if (A.small_t) {
            v[0] = (x <= A.x0) ? A.hL : 0.0;
            v[1] = 0.0;
            return;
        }
        const double xx = x - A.x0;
        if (xx <= A.neg_ts) {
            v[0] = A.hL;
            v[1] = 0.0;
        } else if (xx <= A.two_ts) {
            const double w = A.speed - xx / A.two_t;
            v[0] = A.c49 * (w * w);
            v[1] = A.c827 * (A.speed + xx / A.t) * (w * w);
        } else {
            v[0] = 0.0 + A.htiny;
            v[1] = 0.0;
        }
        return;
    }
    const double dx = x - A.xc, dy = y - A.xc;
    const double bathi = -(A.h0 * (1.0 - (dx * dx + dy * dy) / A.a2));
    if (A.test == 5) {
        const double vv = fmaxF(A.k05 * (2.0 * dx * A.cs + 2.0 * dy * A.sn) - bathi, 0.0);
        v[0] = vv + A.htiny;
        v[1] = -(A.cm * vv);
        v[2] = A.cp * vv;
    } else if (A.test == 11) {
        const double htilde = -bathi - A.A1 - A.cB * dx * A.P1 - A.cB * dy * A.P2;
        const double vv = fmaxF(htilde, 0.0);
        v[0] = vv;
        v[1] = vv * A.gam * A.E2 * A.sn;
        v[2] = vv * A.gam * A.E2 * A.cs;
    } else {
        const double htilde = -bathi - A.cT1 * dy + A.cT2 * dx + A.T3;
        const double vv = fmaxF(htilde, 0.0);
        v[0] = vv;
        v[1] = -(vv * A.delta * A.eta * A.cs);
        v[2] = vv * A.eta * A.sn / A.w;
    }
}

// ---- double-double sums (Knuth's two-sum; the accurate addition of two double-doubles) --------
struct DD {
    double hi, lo;
};
__device__ __forceinline__ DD two_sum(double a, double b) {
    const double s = a + b;
    const double bb = s - a;
    return {s, (a - (s - bb)) + (b - bb)};
}
__device__ __forceinline__ DD fast_two_sum(double a, double b) {
    const double s = a + b;
    return {s, b - (s - a)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    DD s = two_sum(a.hi, b.hi);
    const DD t = two_sum(a.lo, b.lo);
    s.lo = s.lo + t.hi;
    s = fast_two_sum(s.hi, s.lo);
    s.lo = s.lo + t.lo;
    return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ DD dd_add1(DD a, double x) {
    DD s = two_sum(a.hi, x);
    s.lo = s.lo + a.lo;
    return fast_two_sum(s.hi, s.lo);
}

constexpr int kSums = 12;       // per component k: L1 error, L1 norm, L2 error (squared), L2 norm
constexpr int kErrBlock = 256;
constexpr int kErrGrid = 1024;  // fixed partition: the sums do not depend on the launch

struct ErrArgs {
    const int* jj;        // 0-based caller nodes, 3 per element
    const double* rr;     // caller coordinates, 2 per node
    const int* row;       // caller node -> arena row
    const Node* U;
    int64_t me;
};

// The block's NS double-double sums into part[blockIdx.x * NS + q]: a butterfly in a fixed pattern
// within each wave, then the waves in order.
template <int NS>
__device__ __forceinline__ void block_sums(DD (&acc)[NS], DD* __restrict__ part) {
#pragma unroll
    for (int q = 0; q < NS; ++q)
        for (int off = 32; off > 0; off >>= 1) {
            const DD o{__shfl_xor(acc[q].hi, off, 64), __shfl_xor(acc[q].lo, off, 64)};
            acc[q] = dd_add(acc[q], o);
        }
    __shared__ DD wsum[kErrBlock / 64][NS];
    const int wave = (int)(threadIdx.x >> 6);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NS; ++q) wsum[wave][q] = acc[q];
    __syncthreads();
    if (threadIdx.x < NS) {
        DD s = wsum[0][threadIdx.x];
        for (int w = 1; w < kErrBlock / 64; ++w) s = dd_add(s, wsum[w][threadIdx.x]);
        part[(size_t)blockIdx.x * NS + threadIdx.x] = s;
    }
}

// Each thread sums the elements m = tid, tid + nthreads, ...; the wave, then the block, then the grid
// are reduced in a fixed order.
__global__ void __launch_bounds__(kErrBlock) k_errors(ErrArgs a, Anal A, DD* __restrict__ part) {
    const swe_p1::P1Rule R;
    DD acc[kSums];
#pragma unroll
    for (int q = 0; q < kSums; ++q) acc[q] = DD{0.0, 0.0};
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < a.me; m += nthr) {
        int nd[3];
        double f[3][3];
#pragma unroll
        for (int n = 0; n < 3; ++n) nd[n] = a.jj[3 * m + n];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const Node u = a.U[a.row[nd[n]]];
            f[0][n] = u.h;
            f[1][n] = u.qx;
            f[2][n] = u.qy;
        }
        // rj(l,m) of gauss_points_2d_p1 (the P1 Jacobian is the same at the three points)
        double dr[2][2];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int k1 = 0; k1 < 2; ++k1) {
                double s = 0.0;
#pragma unroll
                for (int n = 0; n < 3; ++n) s = s + a.rr[2 * (size_t)nd[n] + k] * R.dd[k1][n];
                dr[k][k1] = s;
            }
        const double rj = (dr[0][0] * dr[1][1] - dr[0][1] * dr[1][0]) * R.pp;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            double rg[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                double s = 0.0;
#pragma unroll
                for (int n = 0; n < 3; ++n) s = s + R.ww[n][l] * a.rr[2 * (size_t)nd[n] + k];
                rg[k] = s;
            }
            double fa[3];
            sol_anal3(A, rg[0], rg[1], fa);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                double fl = 0.0, f0 = 0.0;
#pragma unroll
                for (int n = 0; n < 3; ++n) {
                    fl = fl + f[k][n] * R.ww[n][l];
                    f0 = f0 + 0.0 * R.ww[n][l];
                }
                const double d = fl - fa[k], d0 = f0 - fa[k];
                acc[4 * k + 0] = dd_add1(acc[4 * k + 0], fabs(d) * rj);
                acc[4 * k + 1] = dd_add1(acc[4 * k + 1], fabs(d0) * rj);
                acc[4 * k + 2] = dd_add1(acc[4 * k + 2], d * d * rj);
                acc[4 * k + 3] = dd_add1(acc[4 * k + 3], d0 * d0 * rj);
            }
        }
    }
    block_sums<kSums>(acc, part);
}

// One block: partial sums in block order (lane q of the block reduces sum q).
__global__ void __launch_bounds__(64) k_errors_final(const DD* __restrict__ part, int nblk, double* __restrict__ out) {
    const int q = (int)threadIdx.x;
    if (q >= kSums) return;
    DD s{0.0, 0.0};
    for (int b = 0; b < nblk; ++b) s = dd_add(s, part[(size_t)b * kSums + q]);
    const double v = s.hi + s.lo;
    out[q] = ((q & 3) >= 2) ? sqrt(v) : v;
}

constexpr int kFlowSums = 3;   // ns_l1 of qx, qy and h*sqrt(h)

// The time loop's unit-10 diagnostic (main.f90:94-99): ns_l1 (fem_tn.f90:2119-2153) of qx, qy and the
// nodal h*SQRT(h), the sum over elements and Gauss points of |sum_n f(jj(n,m)) ww(n,l)| rj(l,m).  The
// same partition and reductions as k_errors, in double-double.
__global__ void __launch_bounds__(kErrBlock) k_flow(ErrArgs a, DD* __restrict__ part) {
    const swe_p1::P1Rule R;
    DD acc[kFlowSums];
#pragma unroll
    for (int q = 0; q < kFlowSums; ++q) acc[q] = DD{0.0, 0.0};
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < a.me; m += nthr) {
        int nd[3];
        double f[3][3];
#pragma unroll
        for (int n = 0; n < 3; ++n) nd[n] = a.jj[3 * m + n];
#pragma unroll
        for (int n = 0; n < 3; ++n) {
            const Node u = a.U[a.row[nd[n]]];
            f[0][n] = u.qx;
            f[1][n] = u.qy;
            f[2][n] = u.h * sqrt(u.h);
        }
        double dr[2][2];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int k1 = 0; k1 < 2; ++k1) {
                double s = 0.0;
#pragma unroll
                for (int n = 0; n < 3; ++n) s = s + a.rr[2 * (size_t)nd[n] + k] * R.dd[k1][n];
                dr[k][k1] = s;
            }
        const double rj = (dr[0][0] * dr[1][1] - dr[0][1] * dr[1][0]) * R.pp;
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int k = 0; k < kFlowSums; ++k) {
                double fl = 0.0;
#pragma unroll
                for (int n = 0; n < 3; ++n) fl = fl + f[k][n] * R.ww[n][l];
                acc[k] = dd_add1(acc[k], fabs(fl) * rj);
            }
    }
    block_sums<kFlowSums>(acc, part);
}

__global__ void __launch_bounds__(64) k_flow_final(const DD* __restrict__ part, int nblk, double* __restrict__ out) {
    const int q = (int)threadIdx.x;
    if (q >= kFlowSums) return;
    DD s{0.0, 0.0};
    for (int b = 0; b < nblk; ++b) s = dd_add(s, part[(size_t)b * kFlowSums + q]);
    out[q] = s.hi + s.lo;
}

__global__ void k_node_rows(const int* __restrict__ perm, int64_t r0, int np, int* __restrict__ row) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < np) row[perm[r0 + r]] = (int)(r0 + r);
}

}  // namespace

struct ErrEval {
    int np = 0;
    int64_t me = 0;
    int* jj = nullptr;
    double* rr = nullptr;
    int* row = nullptr;
    DD* part = nullptr;
    double* out = nullptr;
};

void errors_destroy(ErrEval* E) {
    if (!E) return;
    void* bufs[] = {E->jj, E->rr, E->row, E->part, E->out};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    delete E;
}

int errors_create(Pool* P, int k, int me, const int* jj1, const double* rr, ErrEval** out) {
    if (!P || !jj1 || !rr || !out || me <= 0) return fail(SWE_EINVAL, "null or empty argument");
    *out = nullptr;
    if (k < 0 || k >= P->inst_cap || !P->inst[(size_t)k].used) return fail(SWE_EINVAL, "no such instance");
    const InstHost& I = P->inst[(size_t)k];
    const int np = I.np;
    std::vector<int> jj0((size_t)3 * me);
    for (int64_t q = 0; q < 3 * (int64_t)me; ++q) {
        if (jj1[q] < 1 || jj1[q] > np) return fail(SWE_EINVAL, "jj(%lld)=%d out of range", (long long)q + 1, jj1[q]);
        jj0[(size_t)q] = jj1[q] - 1;
    }
    ErrEval* E = new ErrEval();
    E->np = np;
    E->me = me;
    hipStream_t s = P->stream;
    hipError_t e = hipMalloc((void**)&E->jj, sizeof(int) * jj0.size());
    if (e == hipSuccess) e = hipMalloc((void**)&E->rr, sizeof(double) * 2 * (size_t)np);
    if (e == hipSuccess) e = hipMalloc((void**)&E->row, sizeof(int) * (size_t)np);
    if (e == hipSuccess) e = hipMalloc((void**)&E->part, sizeof(DD) * kSums * (size_t)kErrGrid);
    if (e == hipSuccess) e = hipMalloc((void**)&E->out, sizeof(double) * kSums);
    if (e == hipSuccess) e = hipMemcpyAsync(E->jj, jj0.data(), sizeof(int) * jj0.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(E->rr, rr, sizeof(double) * 2 * (size_t)np, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_node_rows, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, (const int*)P->perm,
                           I.slice0 * 64, np, E->row);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        errors_destroy(E);
        return fail(SWE_EHIP, "compute_errors set-up: %s", hipGetErrorString(e));
    }
    *out = E;
    return SWE_OK;
}

int errors_eval(Pool* P, const ErrEval* E, const Node* U, int type_test, double t, double g, double htiny,
                double out[12]) {
    Anal A;
    int rc = make_anal(type_test, t, g, htiny, A);
    if (rc) return rc;
    const int nblk = (int)std::min<int64_t>(kErrGrid, std::max<int64_t>(1, (E->me + kErrBlock - 1) / kErrBlock));
    const ErrArgs a{E->jj, E->rr, E->row, U, E->me};
    hipLaunchKernelGGL(k_errors, dim3((unsigned)nblk), dim3(kErrBlock), 0, P->stream, a, A, E->part);
    hipLaunchKernelGGL(k_errors_final, dim3(1), dim3(64), 0, P->stream, (const DD*)E->part, nblk, E->out);
    SWE_HIPCHK(hipGetLastError());
    SWE_HIPCHK(hipMemcpyAsync(out, E->out, sizeof(double) * kSums, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    return SWE_OK;
}

int errors_flow(Pool* P, const ErrEval* E, const Node* U, double g, double out[4]) {
    const int nblk = (int)std::min<int64_t>(kErrGrid, std::max<int64_t>(1, (E->me + kErrBlock - 1) / kErrBlock));
    const ErrArgs a{E->jj, E->rr, E->row, U, E->me};
    hipLaunchKernelGGL(k_flow, dim3((unsigned)nblk), dim3(kErrBlock), 0, P->stream, a, E->part);
    hipLaunchKernelGGL(k_flow_final, dim3(1), dim3(64), 0, P->stream, (const DD*)E->part, nblk, E->out);
    SWE_HIPCHK(hipGetLastError());
    SWE_HIPCHK(hipMemcpyAsync(out, E->out, sizeof(double) * kFlowSums, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    out[3] = 0.5 * (out[0] + out[1]) / (g * out[2]);   // main.f90:98
    return SWE_OK;
}

}  // namespace swe
