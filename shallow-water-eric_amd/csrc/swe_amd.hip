// swe_amd.hip -- MI355X (gfx950) kernels and C-ABI of the explicit shallow-water step.
//
// One SSP-RK3 step (SOURCES/main.f90:111-135) runs as a short chain of row kernels over the CSR
// graph.  A kernel gives each CSR row to one lane, walks the row's entries in their stored order
// and accumulates every sum sequentially, so each node's result is bitwise the reference's:
//
//   k_compute_dt    COMPUTE_DT + compute_dij  (main.f90:210-218, update.f90:291-326); keeps d_ij
//                   for the first stage and reduces max |d_ii|/m_i per step
//   k_alpha         alpha_limit, first loop   (update.f90:758-781)
//   k_stage_lo      first-order stage: compute_dij/compute_muij/MAX/alpha scaling, smb_2_roundoff
//                   (or smb_2, smb_1), sources, divide_by_lumped, update, RK combination
//                   (update.f90:180-223 and main.f90:121-133)
//   k_fct_lo        FCT stage, pass 1: low-order solution as above plus entropy_residual
//                   (update.f90:825-878) for the high-order viscosity
//   k_fct_rminus    FCT stage, pass 2: compute_fct_matrix + first loop of FCT_positivity
//                   (update.f90:721-748, fct.f90:15-30)
//   k_fct_apply     FCT stage, pass 3: l_ij (fct.f90:31-41), limited viscosities
//                   (update.f90:278-279), apply_viscosity (684-718), update, RK combination
//   k_ev_update     entropy-viscosity-only stage, pass 2 (update.f90:241-262)
//   k_bdy           bdy (main.f90:220-232)
//
// Diagonal entries contribute to the reference's sums only through terms that multiply an exact
// zero (U_i*rat_ii - U_i*rat_ii); those products are skipped, which can change nothing but the
// sign of a zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "swe_amd.h"
#include "swe_kernels.h"

using namespace swe;

namespace {

constexpr int kBlock = 256;
#ifndef SWE_SORT_WINDOW
#define SWE_SORT_WINDOW 64
#endif
constexpr int kSortWindow = SWE_SORT_WINDOW;   // rows sorted by length inside windows of this size

enum SmbKind : int { SMB_ROUNDOFF = 0, SMB_2 = 1, SMB_1 = 2, SMB_GALERKIN = 3 };
enum SrcKind : int { SRC_NONE = 0, SRC_MANNING = 1, SRC_LINEAR = 2, SRC_CORIOLIS = 3 };

// Entry storage: sliced ELL with 64-row slices.  Row r (internal numbering) owns entries
// sptr[r/64] + (r%64) + 64*q for q < len[r], in the row's original CSR order, so a wave's 64
// lanes read 64 consecutive entries per step of q.
struct Mesh {
    const int* __restrict__ len;      // entries of each row (0 for the padding rows of the last slice)
    const int64_t* __restrict__ sptr; // first entry of each slice
    const int* __restrict__ col;      // internal column index of each entry
    const double2* __restrict__ c;    // (cx, cy) of each entry
    const double2* __restrict__ n;    // c_ij/|c_ij| of each entry
    const double* __restrict__ nrm;   // |c_ij| of each entry
    const double* __restrict__ m;     // lumped mass
    const uint16_t* __restrict__ elist; // per slice: slot offsets of its off-diagonal entries
    const int64_t* __restrict__ eoff;   // first elist entry of each slice
    const int* __restrict__ dpos;       // position q of the diagonal in each row
    int np;
    int wmax;                           // longest row (slots per lane in a slice)
};

// LDS of one wave for the compacted lambda phase: per-row one_over_h and SQRT(g*ABS(h)), then
// one double per slot of the slice (slot offset q*64 + lane).
__device__ __forceinline__ double* wave_lds(double* smem, int wmax) {
    return smem + (size_t)(threadIdx.x >> 6) * (size_t)(2 + wmax) * 64;
}

// Phase 0 (one lane per row): row quantities into LDS, and an exact 0.0 in the diagonal slot,
// which is what the reference's row sums add at that position.
__device__ __forceinline__ void lambda_rows(const Mesh& M, const Node& I, bool valid, int i, const Consts& K,
                                            double* lw) {
    const int lane = (int)(threadIdx.x & 63);
    if (valid) {
        lw[lane] = ovh_of(I.h, K.htiny);
        lw[64 + lane] = sqrt(K.g * fabs(I.h));
        lw[128 + 64 * M.dpos[i] + lane] = 0.0;
    }
}

// Phase 1 (entry-parallel): d_ij = |c_ij| lambda_ij of every off-diagonal entry of the wave's
// slice, taken from the slice's compact entry list so that all 64 lanes work on real entries.
// Results go to LDS (and to d_out when given).
__device__ __forceinline__ void lambda_entries(const Mesh& M, const Node* __restrict__ U, const Consts& K,
                                               double* lw, double* __restrict__ d_out, int& err) {
    const int lane = (int)(threadIdx.x & 63);
    const int slice = (int)(((int64_t)xcd_block() * kBlock + (threadIdx.x & ~63u)) >> 6);
    if (slice * 64 >= M.np) return;
    const int64_t e1 = M.eoff[slice + 1];
    const int64_t sb = M.sptr[slice];
    for (int64_t e = M.eoff[slice] + lane; e < e1; e += 64) {
        const int off = M.elist[e];
        const int r = off & 63;
        const int64_t slot = sb + off;
        const Node I = U[slice * 64 + r];
        const Node J = U[M.col[slot]];
        const double d = dij_of(M.n[slot], M.nrm[slot], I, J, lw[r], ovh_of(J.h, K.htiny), lw[64 + r],
                                sqrt(K.g * fabs(J.h)), K, err);
        lw[128 + off] = d;
        if (d_out) d_out[slot] = d;
    }
}

__device__ __forceinline__ int64_t row_base(const Mesh& M, int i) {
    return M.sptr[i >> 6] + (i & 63);
}

__device__ __forceinline__ double read_dt(const double* dtp) { return *dtp; }

__device__ __forceinline__ void raise_flag(int* flags, int f) {
    if (f) atomicOr(flags, f);
}

// RK combination of SOURCES/main.f90:127,133 (stage 0 keeps the Euler output)
__device__ __forceinline__ double rk_combine(int stage, double uo, double ui) {
    if (stage == 1) return (3.0 * uo + ui) / 4.0;
    if (stage == 2) return (uo + 2.0 * ui) / 3.0;
    return ui;
}

// Source terms of one node, dispatched on type_test (update.f90:614-623): friction (425-452),
// my_friction (454-469), coriolis (471-486).  v1, v2 = velocity at U.
template <int SRC>
__device__ __forceinline__ void add_sources(const Node& I, double mi, double v1, double v2, double dt,
                                            const Consts& K, double& r2, double& r3) {
    if constexpr (SRC == SRC_MANNING) {
        const double hs = pow43(fmaxF(I.h, K.htiny));
        const double vel = K.fric * sqrt(v1 * v1 + v2 * v2);
        const double hloc = 1.0 * vel * dt;
        const double den = hs + fmaxF(hs, hloc);
        r2 = r2 - mi * I.qx * vel * 2.0 / den;
        r3 = r3 - mi * I.qy * vel * 2.0 / den;
    } else if constexpr (SRC == SRC_LINEAR) {
        r2 = r2 - mi * 0.2 * I.qx;
        r3 = r3 - mi * 0.2 * I.qy;
    } else if constexpr (SRC == SRC_CORIOLIS) {
        const double f = 2.0;
        r2 = r2 + mi * f * I.qy;
        r3 = r3 - mi * f * I.qx;
    }
}

// =============================================================================================
// COMPUTE_DT: d_ij for every off-diagonal entry, row sums, max |d_ii|/m_i over the grid.
// =============================================================================================
__global__ void __launch_bounds__(kBlock)
k_compute_dt(Mesh M, const Node* __restrict__ U, double* __restrict__ d0, Consts K,
             unsigned long long* __restrict__ vmax, int* __restrict__ flags) {
    extern __shared__ double smem[];
    double* lw = wave_lds(smem, M.wmax);
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    const bool valid = i < M.np;
    const int lane = (int)(threadIdx.x & 63);
    int err = 0;
    Node I{};
    if (valid) I = U[i];
    lambda_rows(M, I, valid, i, K, lw);
    __syncthreads();
    lambda_entries(M, U, K, lw, d0, err);
    __syncthreads();
    double v = 0.0;
    if (valid) {
        // -SUM(dij%aa(ia(i):ia(i+1)-1)) with the diagonal slot holding 0.0, in entry order
        double s = 0.0;
        const int len = M.len[i];
        for (int q = 0; q < len; ++q) s = s + lw[128 + 64 * q + lane];
        v = fabs(-s) / M.m[i];
    }
    raise_flag(flags, err);
    // block max, then one atomic per block on the bit pattern (v >= 0 orders like its bits)
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(b, off, 64);
        b = (o > b) ? o : b;
    }
    __shared__ unsigned long long red[kBlock / 64];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long mx = red[0];
        for (int w = 1; w < kBlock / 64; ++w) mx = (red[w] > mx) ? red[w] : mx;
        atomicMax(vmax, mx);
    }
}

// dt = CFL*1/MAXVAL(...); also resets the accumulator for the next step and logs dt.
__global__ void k_finalize_dt(unsigned long long* vmax, double cfl, double* dtp, double* dt_log, int it) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double mx = __longlong_as_double((long long)*vmax);
        const double dt = cfl * 1.0 / mx;
        *dtp = dt;
        if (dt_log) dt_log[it] = dt;
        *vmax = 0ull;
    }
}

// compute_dij with the full dij%aa (diagonal = -row sum), for swe_compute_dij.
__global__ void __launch_bounds__(kBlock)
k_dij_full(Mesh M, const Node* __restrict__ U, double* __restrict__ dij, Consts K, int* __restrict__ flags) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= M.np) return;
    int err = 0;
    const Node I = U[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double sql = sqrt(K.g * fabs(I.h));
    double s = 0.0;
    int64_t pd = -1;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = M.col[p];
        if (j == i) { pd = p; dij[p] = 0.0; continue; }
        const Node J = U[j];
        const double d = dij_of(M.n[p], M.nrm[p], I, J, ovh_i, ovh_of(J.h, K.htiny), sql, sqrt(K.g * fabs(J.h)), K, err);
        dij[p] = d;
        s = s + d;
    }
    if (pd >= 0) dij[pd] = -s;
    raise_flag(flags, err);
}

// =============================================================================================
// alpha_limit, first loop: alpha_i = (|sum(h_j-h_i)| / sum|h_j-h_i|)**8, 1 if dry.
// =============================================================================================
__global__ void __launch_bounds__(kBlock)
k_alpha(Mesh M, const Node* __restrict__ U, double* __restrict__ alpha, Consts K) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= M.np) return;
    const double hi = U[i].h;
    double a = 1.0;
    if (!(hi <= K.htiny)) {
        const int64_t pb = row_base(M, i);
        const int len = M.len[i];
        // all index loads, then all gathers, then the sums in entry order
        constexpr int kUnroll = 12;
        int js[kUnroll];
        double hs[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) js[q] = (q < len) ? M.col[pb + 64 * q] : i;
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) hs[q] = U[js[q]].h;
        double num = 0.0, denom = 0.0;
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) {
            if (q < len) {
                num = num + hs[q] - hi;
                denom = denom + fabs(hs[q] - hi);
            }
        }
        for (int q = kUnroll; q < len; ++q) {
            const double hj = U[M.col[pb + 64 * (int64_t)q]].h;
            num = num + hj - hi;
            denom = denom + fabs(hj - hi);
        }
        if (denom <= fabs(num)) a = 1.0;
        else a = fabs(num) / (denom);
        a = a * a;
        a = a * a;
        a = a * a;
    }
    alpha[i] = a;
}

// dlam is compute_dij's value for this entry (from the dt pass or from the lambda phase).
template <bool ALPHA>
__device__ __forceinline__ void first_order_dmu(const double2 c, double v1i, double v2i, double v1j, double v2j,
                                                double dlam, double ai, const double* __restrict__ alpha, int j,
                                                double& d, double& mu) {
    d = dlam;
    mu = mu_of(c.x, c.y, v1i, v2i, v1j, v2j);
    d = fmaxF(d, mu);
    if constexpr (ALPHA) {
        const double s = fmaxF(ai, alpha[j]);
        d = d * s;
        mu = mu * s;
    }
}

// =============================================================================================
// First-order stage ('viscous', or 'galerkin'): one pass per stage.
// =============================================================================================
template <int SMB, bool ALPHA, int SRC>
__global__ void __launch_bounds__(kBlock)
k_stage_lo(Mesh M, const Node* __restrict__ U, const Node* __restrict__ Uo, Node* __restrict__ out,
           const double* __restrict__ d0, const double* __restrict__ alpha, const double* __restrict__ dtp,
           Consts K, int stage, int* __restrict__ flags) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    const bool valid = i < M.np;
    int err = 0;
    if (!valid) return;
    const Node I = U[i];
    const double sql = sqrt(K.g * fabs(I.h));
    const double dt = read_dt(dtp);
    const double mi = M.m[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double v1i = I.qx * ovh_i, v2i = I.qy * ovh_i;
    const double ai = ALPHA ? alpha[i] : 1.0;
    const double thr = (SMB == SMB_ROUNDOFF || SMB == SMB_1) ? K.htiny : 0.0;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0, fix = 0.0;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = M.col[p];
        const Node J = U[j];
        const double2 c = M.c[p];
        const double ovh_j = ovh_of(J.h, K.htiny);
        const double v1j = J.qx * ovh_j, v2j = J.qy * ovh_j;
        const double zmax = fmaxF(I.z, J.z);
        const double Hij = fmaxF(0.0, I.h + I.z - zmax);
        const double Hji = fmaxF(0.0, J.h + J.z - zmax);
        const double ratij = (I.h <= thr) ? 0.0 : Hij / I.h;
        const double ratji = (J.h <= thr) ? 0.0 : Hji / J.h;
        double d = 0.0, mu = 0.0;
        const bool off = (j != i);
        if (SMB != SMB_GALERKIN && off)
            first_order_dmu<ALPHA>(c, v1i, v2i, v1j, v2j,
                                   d0 ? d0[p] : dij_of(M.n[p], M.nrm[p], I, J, ovh_i, ovh_j, sql, sqrt(K.g * fabs(J.h)), K, err),
                                   ai, alpha, j, d, mu);
        if constexpr (SMB == SMB_1) {
            // smb_1 (update.f90:378-409)
            double xx = 0.0;
            xx = xx - c.x * (v1j * J.h * ratji + v1i * I.h * ratij);
            xx = xx - c.y * (v2j * J.h * ratji + v2i * I.h * ratij);
            r1 = r1 + xx;
            if (off) r1 = r1 + d * (J.h * ratji - I.h * ratij);
            xx = 0.0;
            xx = xx - c.x * (v1j * J.qx * ratji + v1i * I.qx * ratij);
            xx = xx - c.y * (v2j * J.qx * ratji + v2i * I.qx * ratij);
            r2 = r2 + xx;
            if (off) r2 = r2 + d * (J.qx * ratji - I.qx * ratij);
            xx = 0.0;
            xx = xx - c.x * (v1j * J.qy * ratji + v1i * I.qy * ratij);
            xx = xx - c.y * (v2j * J.qy * ratji + v2i * I.qy * ratij);
            r3 = r3 + xx;
            if (off) r3 = r3 + d * (J.qy * ratji - I.qy * ratij);
            if (off) {
                const double t = 0.5 * K.g * (Hji * Hji - Hij * Hij);
                r2 = r2 - t * c.x;
                r3 = r3 - t * c.y;
            }
        } else {
            if constexpr (SMB == SMB_ROUNDOFF) {
                // smb_2_roundoff (update.f90:566-611): mass equation with fix_roundoff
                double xx = 0.0;
                xx = xx - c.x * (v1j);
                xx = xx - c.y * (v2j);
                if (!off) {
                    fix = fix + xx;
                } else {
                    fix = fix - mu + (d - mu) * (-ratij);
                    r1 = r1 + (xx + mu) * J.h + fmaxF((d - mu) * Hji, 0.0);
                }
            } else {
                // smb_2 (update.f90:656-663), k = 1
                double xx = 0.0;
                xx = xx - c.x * (v1j * J.h);
                xx = xx - c.y * (v2j * J.h);
                r1 = r1 + xx;
                if (SMB != SMB_GALERKIN && off)
                    r1 = r1 + d * (J.h * ratji - I.h * ratij) + mu * (J.h * (1.0 - ratji) - I.h * (1.0 - ratij));
            }
            double xx2 = 0.0;
            xx2 = xx2 - c.x * (v1j * J.qx);
            xx2 = xx2 - c.y * (v2j * J.qx);
            double xx3 = 0.0;
            xx3 = xx3 - c.x * (v1j * J.qy);
            xx3 = xx3 - c.y * (v2j * J.qy);
            r2 = r2 + xx2;
            r3 = r3 + xx3;
            if (SMB != SMB_GALERKIN && off) {
                r2 = r2 + d * (J.qx * ratji - I.qx * ratij) + mu * (J.qx * (1.0 - ratji) - I.qx * (1.0 - ratij));
                r3 = r3 + d * (J.qy * ratji - I.qy * ratij) + mu * (J.qy * (1.0 - ratji) - I.qy * (1.0 - ratij));
            }
            const double topo = K.g * I.h * (J.h + J.z - I.h - I.z);
            r2 = r2 - topo * c.x;
            r3 = r3 - topo * c.y;
        }
    }
    add_sources<SRC>(I, mi, v1i, v2i, dt, K, r2, r3);
    r1 = r1 / mi;
    r2 = r2 / mi;
    r3 = r3 / mi;
    double h;
    if constexpr (SMB == SMB_1) h = I.h + dt * r1;
    else if constexpr (SMB == SMB_ROUNDOFF) h = I.h * (1.0 + dt * fix / mi) + dt * r1;
    else h = fmaxF(I.h + dt * r1, 0.0);
    const double qx = I.qx + dt * r2;
    const double qy = I.qy + dt * r3;
    Node o;
    if (stage == 0) {
        o.h = h; o.qx = qx; o.qy = qy;
    } else {
        const Node P = Uo[i];
        o.h = rk_combine(stage, P.h, h);
        o.qx = rk_combine(stage, P.qx, qx);
        o.qy = rk_combine(stage, P.qy, qy);
    }
    o.z = I.z;
    out[i] = o;
    raise_flag(flags, err);
}

// =============================================================================================
// FCT / entropy-viscosity stage, pass 1: first-order viscosities (kept in dL, muL), the
// low-order solution (smb_2_roundoff, update) and the entropy residual res_i.
// =============================================================================================
template <bool ALPHA, int SRC, bool WRITE_LO>
__global__ void __launch_bounds__(kBlock)
k_fct_lo(Mesh M, const Node* __restrict__ U, Node* __restrict__ ulo, double* __restrict__ dL,
         double* __restrict__ muL, double* __restrict__ res, const double* __restrict__ d0,
         const double* __restrict__ alpha, const double* __restrict__ dtp, Consts K, int roundoff,
         int* __restrict__ flags) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    const bool valid = i < M.np;
    int err = 0;
    if (!valid) return;
    const Node I = U[i];
    const double sql = sqrt(K.g * fabs(I.h));
    const double dt = read_dt(dtp);
    const double mi = M.m[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double v1i = I.qx * ovh_i, v2i = I.qy * ovh_i;
    const double ai = ALPHA ? alpha[i] : 1.0;
    const double thr_lo = roundoff ? K.htiny : 0.0;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0, fix = 0.0;   // smb_2_roundoff (or smb_2)
    double e1 = 0.0, e2 = 0.0, e3 = 0.0;              // smb_2 of entropy_residual
    double emax = 0.0, emin = 0.0;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = M.col[p];
        const Node J = U[j];
        const double2 c = M.c[p];
        const double ovh_j = ovh_of(J.h, K.htiny);
        const double v1j = J.qx * ovh_j, v2j = J.qy * ovh_j;
        const double zmax = fmaxF(I.z, J.z);
        const double Hij = fmaxF(0.0, I.h + I.z - zmax);
        const double Hji = fmaxF(0.0, J.h + J.z - zmax);
        const double qij = Hij / I.h, qji = Hji / J.h;
        const double ratij0 = (I.h <= 0.0) ? 0.0 : qij;
        const double ratji0 = (J.h <= 0.0) ? 0.0 : qji;
        const double ratijr = (I.h <= thr_lo) ? 0.0 : qij;
        const double ratjir = (J.h <= thr_lo) ? 0.0 : qji;
        const bool off = (j != i);
        double d = 0.0, mu = 0.0;
        if (off) {
            first_order_dmu<ALPHA>(c, v1i, v2i, v1j, v2j,
                                   d0 ? d0[p] : dij_of(M.n[p], M.nrm[p], I, J, ovh_i, ovh_j, sql, sqrt(K.g * fabs(J.h)), K, err),
                                   ai, alpha, j, d, mu);
            dL[p] = d;
            muL[p] = mu;
        }
        double xm = 0.0;
        xm = xm - c.x * (v1j);
        xm = xm - c.y * (v2j);
        double xe = 0.0;
        xe = xe - c.x * (v1j * J.h);
        xe = xe - c.y * (v2j * J.h);
        double xx2 = 0.0;
        xx2 = xx2 - c.x * (v1j * J.qx);
        xx2 = xx2 - c.y * (v2j * J.qx);
        double xx3 = 0.0;
        xx3 = xx3 - c.x * (v1j * J.qy);
        xx3 = xx3 - c.y * (v2j * J.qy);
        const double topo = K.g * I.h * (J.h + J.z - I.h - I.z);
        if (WRITE_LO) {
            if (roundoff) {
                if (!off) fix = fix + xm;
                else {
                    fix = fix - mu + (d - mu) * (-ratijr);
                    r1 = r1 + (xm + mu) * J.h + fmaxF((d - mu) * Hji, 0.0);
                }
            } else {
                r1 = r1 + xe;
                if (off) r1 = r1 + d * (J.h * ratjir - I.h * ratijr) + mu * (J.h * (1.0 - ratjir) - I.h * (1.0 - ratijr));
            }
            r2 = r2 + xx2;
            r3 = r3 + xx3;
            if (off) {
                r2 = r2 + d * (J.qx * ratjir - I.qx * ratijr) + mu * (J.qx * (1.0 - ratjir) - I.qx * (1.0 - ratijr));
                r3 = r3 + d * (J.qy * ratjir - I.qy * ratijr) + mu * (J.qy * (1.0 - ratjir) - I.qy * (1.0 - ratijr));
            }
            r2 = r2 - topo * c.x;
            r3 = r3 - topo * c.y;
        }
        e1 = e1 + xe;
        e2 = e2 + xx2;
        e3 = e3 + xx3;
        if (off) {
            e1 = e1 + d * (J.h * ratji0 - I.h * ratij0) + mu * (J.h * (1.0 - ratji0) - I.h * (1.0 - ratij0));
            e2 = e2 + d * (J.qx * ratji0 - I.qx * ratij0) + mu * (J.qx * (1.0 - ratji0) - I.qx * (1.0 - ratij0));
            e3 = e3 + d * (J.qy * ratji0 - I.qy * ratij0) + mu * (J.qy * (1.0 - ratji0) - I.qy * (1.0 - ratij0));
        }
        e2 = e2 - topo * c.x;
        e3 = e3 - topo * c.y;
        // entropy of the neighbour for maxmin (update.f90:924-935)
        double Ej = K.g * (J.h * J.h);
        Ej = Ej + v1j * J.qx;
        Ej = Ej + v2j * J.qy;
        Ej = 0.5 * Ej;
        if (q == 0) { emax = Ej; emin = Ej; }
        else {
            emax = (Ej > emax) ? Ej : emax;
            emin = (Ej < emin) ? Ej : emin;
        }
    }
    add_sources<SRC>(I, mi, v1i, v2i, dt, K, e2, e3);
    if (WRITE_LO) {
        add_sources<SRC>(I, mi, v1i, v2i, dt, K, r2, r3);
        r1 = r1 / mi;
        r2 = r2 / mi;
        r3 = r3 / mi;
        Node o;
        if (roundoff) o.h = I.h * (1.0 + dt * fix / mi) + dt * r1;
        else o.h = fmaxF(I.h + dt * r1, 0.0);
        o.qx = I.qx + dt * r2;
        o.qy = I.qy + dt * r3;
        o.z = I.z;
        ulo[i] = o;
        if (K.check_h && o.h < 0.0) err |= FLAG_NEG_H;
    }
    // entropy_residual (update.f90:838-868)
    e1 = e1 / mi;
    e2 = e2 / mi;
    e3 = e3 / mi;
    const double ht = I.h + dt * e1;
    const double qxt = I.qx + dt * e2;
    const double qyt = I.qy + dt * e3;
    const double ovt = ovh_of(ht, K.htiny);
    const double vn1 = qxt * ovt, vn2 = qyt * ovt;
    double E = K.g * (I.h * I.h);
    double En = K.g * (ht * ht);
    E = E + v1i * I.qx;
    En = En + vn1 * qxt;
    E = E + v2i * I.qy;
    En = En + vn2 * qyt;
    E = 0.5 * E;
    En = 0.5 * En;
    double Ep0 = K.g * I.h;
    Ep0 = Ep0 - 0.5 * (v1i * v1i);
    Ep0 = Ep0 - 0.5 * (v2i * v2i);
    double r = mi * (En - E) / dt;
    r = r - mi * e1 * Ep0;
    r = r - mi * e2 * v1i;
    r = r - mi * e3 * v2i;
    const double rescale = fabs(emax - emin) / 2.0 + K.ent_cst;
    res[i] = fabs(r) / rescale;
    raise_flag(flags, err);
}

// Antidiffusive flux f_ij of compute_fct_matrix (update.f90:732-745) for an off-diagonal entry,
// with the high-order viscosities MIN(dL, resij), MIN(muL, resij) of update.f90:233-239.
__device__ __forceinline__ double fct_flux(const Node& I, const Node& J, double dLp, double muLp,
                                           double rij, double& dH, double& muH, double& ratij, double& ratji) {
    dH = fminF(dLp, 1.0 * rij);
    muH = fminF(muLp, 1.0 * rij);
    const double zmax = fmaxF(I.z, J.z);
    const double Hij = fmaxF(0.0, I.h + I.z - zmax);
    const double Hji = fmaxF(0.0, J.h + J.z - zmax);
    ratij = (I.h <= 0.0) ? 0.0 : Hij / I.h;
    ratji = (J.h <= 0.0) ? 0.0 : Hji / J.h;
    return (dH - dLp) * (J.h * ratji - I.h * ratij) + (muH - muLp) * (J.h * (1.0 - ratji) - I.h * (1.0 - ratij));
}

// =============================================================================================
// FCT stage, pass 2: P-, Q-, R- of FCT_positivity (fct.f90:15-30).
// =============================================================================================
__global__ void __launch_bounds__(kBlock)
k_fct_rminus(Mesh M, const Node* __restrict__ U, const Node* __restrict__ ulo,
             const double* __restrict__ dL, const double* __restrict__ muL, const double* __restrict__ res,
             double* __restrict__ Rm, const double* __restrict__ dtp, Consts K) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= M.np) return;
    const double dt = read_dt(dtp);
    const Node I = U[i];
    const double resi = res[i];
    const double scaled = (M.m[i] / dt) * ulo[i].h;
    const double Q = 0.0 - scaled;
    double P = -1.0e-15;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = M.col[p];
        if (j == i) continue;    // f_ii is a zero, never < 0
        const Node J = U[j];
        double dH, muH, rij_, rji_;
        const double f = fct_flux(I, J, dL[p], muL[p], fmaxF(resi, res[j]), dH, muH, rij_, rji_);
        if (f < 0.0) P = P + f;
    }
    double R = fminF(Q / P, 1.0);
    if (scaled <= K.htiny) R = 0.0;
    Rm[i] = R;
}

// =============================================================================================
// FCT stage, pass 3: l_ij, limited viscosities, apply_viscosity, update, RK combination.
// =============================================================================================
__global__ void __launch_bounds__(kBlock)
k_fct_apply(Mesh M, const Node* __restrict__ U, const Node* __restrict__ Uo, const Node* __restrict__ ulo,
            Node* __restrict__ out,
            const double* __restrict__ dL, const double* __restrict__ muL, const double* __restrict__ res,
            const double* __restrict__ Rm, const double* __restrict__ dtp, Consts K, int stage,
            int* __restrict__ flags) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= M.np) return;
    const double dt = read_dt(dtp);
    const Node I = U[i];
    const double resi = res[i];
    const double Ri = Rm[i];
    double x1 = 0.0, x2 = 0.0, x3 = 0.0;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = M.col[p];
        if (j == i) continue;    // contributes d*(0) + mu*(0)
        const Node J = U[j];
        const double dLp = dL[p], muLp = muL[p];
        double dH, muH, ratij, ratji;
        const double f = fct_flux(I, J, dLp, muLp, fmaxF(resi, res[j]), dH, muH, ratij, ratji);
        const double l = (f >= 0.0) ? Rm[j] : Ri;
        const double d = (dH - dLp) * l;
        const double mu = (muH - muLp) * l;
        x1 = x1 + d * (J.h * ratji - I.h * ratij) + mu * (J.h * (1.0 - ratji) - I.h * (1.0 - ratij));
        x2 = x2 + d * (J.qx * ratji - I.qx * ratij) + mu * (J.qx * (1.0 - ratji) - I.qx * (1.0 - ratij));
        x3 = x3 + d * (J.qy * ratji - I.qy * ratij) + mu * (J.qy * (1.0 - ratji) - I.qy * (1.0 - ratij));
    }
    const double mi = M.m[i];
    const Node L = ulo[i];      // low-order solution of k_fct_lo
    const double h = L.h + dt * (x1 / mi);
    const double qx = L.qx + dt * (x2 / mi);
    const double qy = L.qy + dt * (x3 / mi);
    int err = (K.check_h && h < 0.0) ? FLAG_NEG_H : 0;
    Node o;
    if (stage == 0) {
        o.h = h; o.qx = qx; o.qy = qy;
    } else {
        const Node P = Uo[i];
        o.h = rk_combine(stage, P.h, h);
        o.qx = rk_combine(stage, P.qx, qx);
        o.qy = rk_combine(stage, P.qy, qy);
    }
    o.z = I.z;
    out[i] = o;
    raise_flag(flags, err);
}

// =============================================================================================
// Entropy-viscosity-only stage, pass 2 (update.f90:241-262).
// =============================================================================================
template <int SRC>
__global__ void __launch_bounds__(kBlock)
k_ev_update(Mesh M, const Node* __restrict__ U, const Node* __restrict__ Uo, Node* __restrict__ out,
            const double* __restrict__ dL, const double* __restrict__ muL, const double* __restrict__ res,
            const double* __restrict__ dtp, Consts K, int roundoff, int stage, int* __restrict__ flags) {
    const int i = xcd_block() * kBlock + (int)threadIdx.x;
    if (i >= M.np) return;
    const double dt = read_dt(dtp);
    const Node I = U[i];
    const double mi = M.m[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double v1i = I.qx * ovh_i, v2i = I.qy * ovh_i;
    const double resi = res[i];
    const double thr = roundoff ? K.htiny : 0.0;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0, fix = 0.0;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = M.col[p];
        const Node J = U[j];
        const double2 c = M.c[p];
        const double ovh_j = ovh_of(J.h, K.htiny);
        const double v1j = J.qx * ovh_j, v2j = J.qy * ovh_j;
        const double zmax = fmaxF(I.z, J.z);
        const double Hij = fmaxF(0.0, I.h + I.z - zmax);
        const double Hji = fmaxF(0.0, J.h + J.z - zmax);
        const double ratij = (I.h <= thr) ? 0.0 : Hij / I.h;
        const double ratji = (J.h <= thr) ? 0.0 : Hji / J.h;
        const bool off = (j != i);
        double d = 0.0, mu = 0.0;
        if (off) {
            const double rij = fmaxF(resi, res[j]);
            d = fminF(dL[p], 1.0 * rij);
            mu = fminF(muL[p], 1.0 * rij);
        }
        if (roundoff) {
            double xm = 0.0;
            xm = xm - c.x * (v1j);
            xm = xm - c.y * (v2j);
            if (!off) fix = fix + xm;
            else {
                fix = fix - mu + (d - mu) * (-ratij);
                r1 = r1 + (xm + mu) * J.h + fmaxF((d - mu) * Hji, 0.0);
            }
        } else {
            double xe = 0.0;
            xe = xe - c.x * (v1j * J.h);
            xe = xe - c.y * (v2j * J.h);
            r1 = r1 + xe;
            if (off) r1 = r1 + d * (J.h * ratji - I.h * ratij) + mu * (J.h * (1.0 - ratji) - I.h * (1.0 - ratij));
        }
        double xx2 = 0.0;
        xx2 = xx2 - c.x * (v1j * J.qx);
        xx2 = xx2 - c.y * (v2j * J.qx);
        double xx3 = 0.0;
        xx3 = xx3 - c.x * (v1j * J.qy);
        xx3 = xx3 - c.y * (v2j * J.qy);
        r2 = r2 + xx2;
        r3 = r3 + xx3;
        if (off) {
            r2 = r2 + d * (J.qx * ratji - I.qx * ratij) + mu * (J.qx * (1.0 - ratji) - I.qx * (1.0 - ratij));
            r3 = r3 + d * (J.qy * ratji - I.qy * ratij) + mu * (J.qy * (1.0 - ratji) - I.qy * (1.0 - ratij));
        }
        const double topo = K.g * I.h * (J.h + J.z - I.h - I.z);
        r2 = r2 - topo * c.x;
        r3 = r3 - topo * c.y;
    }
    add_sources<SRC>(I, mi, v1i, v2i, dt, K, r2, r3);
    r1 = r1 / mi;
    r2 = r2 / mi;
    r3 = r3 / mi;
    double h;
    if (roundoff) h = I.h * (1.0 + dt * fix / mi) + dt * r1;
    else h = fmaxF(I.h + dt * r1, 0.0);
    const double qx = I.qx + dt * r2;
    const double qy = I.qy + dt * r3;
    int err = (K.check_h && h < 0.0) ? FLAG_NEG_H : 0;
    Node o;
    if (stage == 0) {
        o.h = h; o.qx = qx; o.qy = qy;
    } else {
        const Node P = Uo[i];
        o.h = rk_combine(stage, P.h, h);
        o.qx = rk_combine(stage, P.qx, qx);
        o.qy = rk_combine(stage, P.qy, qy);
    }
    o.z = I.z;
    out[i] = o;
    raise_flag(flags, err);
}

// =============================================================================================
// bdy and layout conversion
// =============================================================================================
__global__ void k_bdy(Node* __restrict__ u, const int* __restrict__ hl, int nh, const int* __restrict__ uxl,
                      int nux, const int* __restrict__ uyl, int nuy, int kind, double level, double amp_sin,
                      const double* __restrict__ hv, const double* __restrict__ uxv, const double* __restrict__ uyv) {
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q < nh) {
        const int n = hl[q];
        if (kind == SWE_BC_INFLOW) u[n].h = fmaxF(level - u[n].z + amp_sin, 0.0);
        else if (kind == SWE_BC_VALUES) u[n].h = hv[q];
    }
    if (q < nux) u[uxl[q]].qx = (kind == SWE_BC_VALUES) ? uxv[q] : 0.0;
    if (q < nuy) u[uyl[q]].qy = (kind == SWE_BC_VALUES) ? uyv[q] : 0.0;
}

// un(3,np) in the caller's numbering <-> internal Node records; perm maps internal -> caller.
__global__ void k_pack(const double* __restrict__ un3, const double* __restrict__ z, const int* __restrict__ perm,
                       Node* __restrict__ u, int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np) return;
    const size_t i = (size_t)perm[r];
    Node o;
    o.h = un3[3 * i];
    o.qx = un3[3 * i + 1];
    o.qy = un3[3 * i + 2];
    o.z = z[r];
    u[r] = o;
}

__global__ void k_unpack(const Node* __restrict__ u, const int* __restrict__ perm, double* __restrict__ un3, int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np) return;
    const size_t i = (size_t)perm[r];
    const Node o = u[r];
    un3[3 * i] = o.h;
    un3[3 * i + 1] = o.qx;
    un3[3 * i + 2] = o.qy;
}

// node array in the caller's numbering -> internal numbering
__global__ void k_perm_nodes(const double* __restrict__ src, const int* __restrict__ perm, double* __restrict__ dst, int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < np) dst[r] = src[perm[r]];
}

// entry array in CSR order <-> sliced layout, through the CSR -> slot map
__global__ void k_csr_to_sell(const double* __restrict__ src, const int64_t* __restrict__ map, double* __restrict__ dst, int64_t nnz) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) dst[map[p]] = src[p];
}

__global__ void k_sell_to_csr(const double* __restrict__ src, const int64_t* __restrict__ map, double* __restrict__ dst, int64_t nnz) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) dst[p] = src[map[p]];
}

// FCT_positivity on arbitrary fluxes (swe_fct_positivity)
__global__ void k_fctpos_r(Mesh M, const double* __restrict__ f, const double* __restrict__ us,
                           double* __restrict__ Rm, double htiny) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= M.np) return;
    const double Q = 0.0 - us[i];
    double P = -1.0e-15;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const double fij = f[p];
        if (fij < 0.0) P = P + fij;
    }
    double R = fminF(Q / P, 1.0);
    if (us[i] <= htiny) R = 0.0;
    Rm[i] = R;
}

__global__ void k_fctpos_l(Mesh M, const double* __restrict__ f, const double* __restrict__ Rm,
                           double* __restrict__ lij) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= M.np) return;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        lij[p] = (f[p] >= 0.0) ? Rm[M.col[p]] : Rm[i];
    }
}

// FCT_positivity on a caller's 1-based CSR matrix (swe_fct_positivity_csr)
__global__ void k_fctpos_csr_r(int np, const int* __restrict__ ia, const double* __restrict__ f,
                               const double* __restrict__ us, double* __restrict__ Rm, double htiny) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const double Q = 0.0 - us[i];
    double P = -1.0e-15;
    for (int p = ia[i] - 1; p < ia[i + 1] - 1; ++p) {
        const double fij = f[p];
        if (fij < 0.0) P = P + fij;
    }
    double R = fminF(Q / P, 1.0);
    if (us[i] <= htiny) R = 0.0;
    Rm[i] = R;
}

__global__ void k_fctpos_csr_l(int np, const int* __restrict__ ia, const int* __restrict__ ja,
                               const double* __restrict__ f, const double* __restrict__ Rm, double* __restrict__ lij) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    for (int p = ia[i] - 1; p < ia[i + 1] - 1; ++p) lij[p] = (f[p] >= 0.0) ? Rm[ja[p] - 1] : Rm[i];
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                                \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return fail(SWE_EHIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

template <class T>
int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(SWE_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T), hipGetErrorString(e));
    return SWE_OK;
}

struct TimedKernel {
    std::vector<hipEvent_t> ev;  // start/stop pairs
    int used = 0;
};

}  // namespace

struct swe_handle {
    int np = 0, nnz = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    // mesh in the internal numbering and the sliced entry layout
    int np_pad = 0;                 // np rounded up to whole 64-row slices
    int64_t nslot = 0;              // entry slots of the sliced layout (>= nnz)
    int* len = nullptr;
    int64_t* sptr = nullptr;
    int* perm = nullptr;            // internal row -> caller's node
    int64_t* csr2slot = nullptr;    // caller's CSR entry -> slot
    int* col = nullptr;
    double2* c = nullptr;
    double2* nvec = nullptr;        // c_ij/|c_ij|
    double* nrm = nullptr;          // |c_ij|
    uint16_t* elist = nullptr;      // per slice: slot offsets of its off-diagonal entries
    int64_t* eoff = nullptr;
    int* dpos = nullptr;            // diagonal position in each row
    int wmax = 0;                   // longest row
    size_t lds_bytes = 0;           // dynamic LDS of the kernels with a lambda phase
    double* m = nullptr;
    double* z = nullptr;
    // states: S[0] holds U^n between steps
    Node* S[3] = {nullptr, nullptr, nullptr};
    Node* scratch = nullptr;
    Node* ulo = nullptr;            // low-order solution of an FCT stage
    double* un3 = nullptr;          // (3,np) staging for the host-array entry points
    // per-entry and per-node work arrays
    double *d0 = nullptr, *dL = nullptr, *muL = nullptr, *ework = nullptr;
    double *alpha = nullptr, *res = nullptr, *Rm = nullptr, *nwork = nullptr;
    // scalars
    unsigned long long* vmax = nullptr;
    double* dtp = nullptr;
    double* dt_log = nullptr;
    int dt_log_cap = 0;
    int* flags = nullptr;
    double* host_pinned = nullptr;  // [0] dt, [1..] misc
    double time = 0.0;              // simulation time (inputs%time)
    int pending = 0;                // queued steps whose dt is not yet folded into `time`
    // boundary conditions
    int nh = 0, nux = 0, nuy = 0;
    int *hl = nullptr, *uxl = nullptr, *uyl = nullptr;
    double *hv = nullptr, *uxv = nullptr, *uyv = nullptr;
    int bc_kind = SWE_BC_NONE;
    double bc_amp = 0.0, bc_period = 1.0, bc_level = 0.0;
    // parameters
    swe_params prm{};
    Consts K{};
    bool state_set = false;
    // timing
    bool timing = false;
    TimedKernel tk;
    std::string timed_name;
};

namespace {

Mesh mesh_of(const swe_handle* h) {
    return Mesh{h->len, h->sptr, h->col, h->c, h->nvec, h->nrm, h->m, h->elist, h->eoff, h->dpos, h->np, h->wmax};
}

int make_consts(const swe_params& p, Consts& K) {
    if (!(p.gravity > 0.0)) return fail(SWE_EINVAL, "gravity must be > 0");
    if (p.viscosity_type < SWE_GALERKIN || p.viscosity_type > SWE_FCT)
        return fail(SWE_EBADOPT, "BUG in euler: unknown viscosity_type %d", p.viscosity_type);
    if (p.viscous_type != 1 && p.viscous_type != 2)
        return fail(SWE_EBADOPT, "BUG in euler, viscous_type %d", p.viscous_type);
    if (p.viscosity_type == SWE_ENTROPY_VISC && p.viscous_type == 1)
        return fail(SWE_EBADOPT, "Bug: entropy viscosity programmed only with type 2");
    if (p.viscosity_type == SWE_FCT && p.viscous_type != 2)
        return fail(SWE_EBADOPT, "BUG: FCT programmed with type2 only");
    K.g = p.gravity;
    K.htiny = p.htiny;
    K.cfl = p.cfl;
    K.max_water_h = p.max_water_h;
    K.mannings = p.mannings;
    const double t = 2.0 * std::sqrt(2.0) - 1.0;
    K.x0 = t * t;
    K.a_c2 = 1 / (2.0 * std::sqrt(2.0));
    K.two_sqrt_g = 2 * std::sqrt(p.gravity);
    K.sqrt_2_g = std::sqrt(2 / p.gravity);
    K.g_half = p.gravity / 2;
    K.g16 = 16 * p.gravity;
    K.fric = p.gravity * (p.mannings * p.mannings);
    K.ent_cst = 1.0e-10 * p.gravity * (p.max_water_h * p.max_water_h);
    K.type_test = p.type_test;
    K.check_h = (p.type_test >= 1 && p.type_test <= 13) ? 1 : 0;
    return SWE_OK;
}

int src_kind(int type_test) {
    if (type_test == 9 || type_test == 10) return SRC_MANNING;
    if (type_test == 11) return SRC_LINEAR;
    if (type_test == 12) return SRC_CORIOLIS;
    return SRC_NONE;
}

inline unsigned grid_of(int n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void timed_begin(swe_handle* h, const char* name) {
    if (!h->timing || h->timed_name != name) return;
    TimedKernel& T = h->tk;
    if ((size_t)(2 * T.used + 2) > T.ev.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            T.ev.push_back(e);
        }
    }
    (void)hipEventRecord(T.ev[(size_t)2 * T.used], h->stream);
}

void timed_end(swe_handle* h, const char* name) {
    if (!h->timing || h->timed_name != name) return;
    TimedKernel& T = h->tk;
    (void)hipEventRecord(T.ev[(size_t)2 * T.used + 1], h->stream);
    T.used++;
}

template <int SMB, bool ALPHA>
void launch_lo_src(swe_handle* h, int src, const Node* U, const Node* Uo, Node* out, const double* d0, int stage) {
    const Mesh M = mesh_of(h);
    const unsigned g = grid_of(h->np);
#define SWE_LO(S) hipLaunchKernelGGL((k_stage_lo<SMB, ALPHA, S>), dim3(g), dim3(kBlock), 0, h->stream, M, U, Uo, out, d0, h->alpha, h->dtp, h->K, stage, h->flags)
    switch (src) {
        case SRC_MANNING: SWE_LO(SRC_MANNING); break;
        case SRC_LINEAR: SWE_LO(SRC_LINEAR); break;
        case SRC_CORIOLIS: SWE_LO(SRC_CORIOLIS); break;
        default: SWE_LO(SRC_NONE); break;
    }
#undef SWE_LO
}

template <bool ALPHA, bool WRITE_LO>
void launch_fct_lo_src(swe_handle* h, int src, const Node* U, Node* ulo, const double* d0) {
    const Mesh M = mesh_of(h);
    const unsigned g = grid_of(h->np);
    const int ro = h->prm.roundoff_fix ? 1 : 0;
#define SWE_FL(S) hipLaunchKernelGGL((k_fct_lo<ALPHA, S, WRITE_LO>), dim3(g), dim3(kBlock), 0, h->stream, M, U, ulo, h->dL, h->muL, h->res, d0, h->alpha, h->dtp, h->K, ro, h->flags)
    switch (src) {
        case SRC_MANNING: SWE_FL(SRC_MANNING); break;
        case SRC_LINEAR: SWE_FL(SRC_LINEAR); break;
        case SRC_CORIOLIS: SWE_FL(SRC_CORIOLIS); break;
        default: SWE_FL(SRC_NONE); break;
    }
#undef SWE_FL
}

// One Euler stage from U to out (out != U).  d0 (may be null) holds compute_dij at U.
// stage: 0 keeps the Euler output, 1 and 2 apply the RK combinations with Uo.
int euler_stage(swe_handle* h, const Node* U, const Node* Uo, Node* out, const double* d0, int stage) {
    const swe_params& p = h->prm;
    const Mesh M = mesh_of(h);
    const unsigned g = grid_of(h->np);
    const int src = src_kind(p.type_test);
    const bool alpha = p.alpha_limit != 0 && p.viscosity_type != SWE_GALERKIN;
    if (alpha) hipLaunchKernelGGL(k_alpha, dim3(g), dim3(kBlock), 0, h->stream, M, U, h->alpha, h->K);
    if (p.viscosity_type == SWE_GALERKIN) {
        launch_lo_src<SMB_GALERKIN, false>(h, src, U, Uo, out, nullptr, stage);
    } else if (p.viscosity_type == SWE_VISCOUS) {
        const int smb = (p.viscous_type == 1) ? SMB_1 : (p.roundoff_fix ? SMB_ROUNDOFF : SMB_2);
        timed_begin(h, "k_stage_lo");
        if (smb == SMB_1) {
            if (alpha) launch_lo_src<SMB_1, true>(h, src, U, Uo, out, d0, stage);
            else launch_lo_src<SMB_1, false>(h, src, U, Uo, out, d0, stage);
        } else if (smb == SMB_ROUNDOFF) {
            if (alpha) launch_lo_src<SMB_ROUNDOFF, true>(h, src, U, Uo, out, d0, stage);
            else launch_lo_src<SMB_ROUNDOFF, false>(h, src, U, Uo, out, d0, stage);
        } else {
            if (alpha) launch_lo_src<SMB_2, true>(h, src, U, Uo, out, d0, stage);
            else launch_lo_src<SMB_2, false>(h, src, U, Uo, out, d0, stage);
        }
        timed_end(h, "k_stage_lo");
    } else if (p.viscosity_type == SWE_FCT) {
        timed_begin(h, "k_fct_lo");
        if (alpha) launch_fct_lo_src<true, true>(h, src, U, h->ulo, d0);
        else launch_fct_lo_src<false, true>(h, src, U, h->ulo, d0);
        timed_end(h, "k_fct_lo");
        timed_begin(h, "k_fct_rminus");
        hipLaunchKernelGGL(k_fct_rminus, dim3(g), dim3(kBlock), 0, h->stream, M, U, (const Node*)h->ulo,
                           (const double*)h->dL, (const double*)h->muL, (const double*)h->res, h->Rm,
                           (const double*)h->dtp, h->K);
        timed_end(h, "k_fct_rminus");
        timed_begin(h, "k_fct_apply");
        hipLaunchKernelGGL(k_fct_apply, dim3(g), dim3(kBlock), 0, h->stream, M, U, Uo, (const Node*)h->ulo, out,
                           (const double*)h->dL, (const double*)h->muL, (const double*)h->res,
                           (const double*)h->Rm, (const double*)h->dtp, h->K, stage, h->flags);
        timed_end(h, "k_fct_apply");
    } else {  // entropy_visc
        if (alpha) launch_fct_lo_src<true, false>(h, src, U, out, d0);
        else launch_fct_lo_src<false, false>(h, src, U, out, d0);
        const int ro = p.roundoff_fix ? 1 : 0;
#define SWE_EV(S) hipLaunchKernelGGL((k_ev_update<S>), dim3(g), dim3(kBlock), 0, h->stream, M, U, Uo, out, (const double*)h->dL, (const double*)h->muL, (const double*)h->res, (const double*)h->dtp, h->K, ro, stage, h->flags)
        switch (src) {
            case SRC_MANNING: SWE_EV(SRC_MANNING); break;
            case SRC_LINEAR: SWE_EV(SRC_LINEAR); break;
            case SRC_CORIOLIS: SWE_EV(SRC_CORIOLIS); break;
            default: SWE_EV(SRC_NONE); break;
        }
#undef SWE_EV
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SWE_EHIP, "kernel launch failed: %s", hipGetErrorString(e));
    return SWE_OK;
}

bool uses_lambda(const swe_params& p) { return p.viscosity_type != SWE_GALERKIN; }

void launch_bdy(swe_handle* h, Node* u, double t) {
    const int n = std::max(h->nh, std::max(h->nux, h->nuy));
    if (n == 0 || h->bc_kind == SWE_BC_NONE) return;
    double amp_sin = 0.0;
    if (h->bc_kind == SWE_BC_INFLOW) {
        const double pi = std::acos(-1.0);
        amp_sin = h->bc_amp * std::sin(2.0 * pi * t / h->bc_period);
    }
    hipLaunchKernelGGL(k_bdy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, u, h->hl, h->nh,
                       h->uxl, h->nux, h->uyl, h->nuy, h->bc_kind, h->bc_level, amp_sin, h->hv, h->uxv, h->uyv);
}

int check_flags(swe_handle* h) {
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, h->flags, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (f) {
        HIPCHK(hipMemsetAsync(h->flags, 0, sizeof(int), h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (f & FLAG_LAMBDA) return fail(SWE_ELAMBDA, "BUG in compute_lambda_vacc (Delta < 0)");
        if (f & FLAG_NEG_H) return fail(SWE_ENEG_H, "Min h<0, STOP");
    }
    return SWE_OK;
}

// Fold the dt of queued steps into the handle's time: time = to + dt per step, in order
// (SOURCES/main.f90:135).
int settle_time(swe_handle* h) {
    if (h->pending == 0) return SWE_OK;
    std::vector<double> dts((size_t)h->pending);
    HIPCHK(hipMemcpyAsync(dts.data(), h->dt_log, sizeof(double) * dts.size(), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (double dt : dts) {
        const double to = h->time;
        h->time = to + dt;
    }
    h->pending = 0;
    return SWE_OK;
}

int ensure_dt_log(swe_handle* h, int n) {
    if (h->pending + n <= h->dt_log_cap) return SWE_OK;
    int rc = settle_time(h);
    if (rc) return rc;
    if (n <= h->dt_log_cap) return SWE_OK;
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->dt_log) (void)hipFree(h->dt_log);
    h->dt_log = nullptr;
    h->dt_log_cap = 0;
    const int cap = std::max(n, 4096);
    if ((rc = dalloc(&h->dt_log, (size_t)cap))) return rc;
    h->dt_log_cap = cap;
    return SWE_OK;
}

int upload_state(swe_handle* h, const double* un, Node* dst) {
    HIPCHK(hipMemcpyAsync(h->un3, un, sizeof(double) * 3 * (size_t)h->np, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_pack, dim3(grid_of(h->np)), dim3(kBlock), 0, h->stream, (const double*)h->un3,
                       (const double*)h->z, (const int*)h->perm, dst, h->np);
    return SWE_OK;
}

int download_state(swe_handle* h, const Node* src, double* un) {
    hipLaunchKernelGGL(k_unpack, dim3(grid_of(h->np)), dim3(kBlock), 0, h->stream, src, (const int*)h->perm, h->un3, h->np);
    HIPCHK(hipMemcpyAsync(un, h->un3, sizeof(double) * 3 * (size_t)h->np, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SWE_OK;
}

// Queue `nsteps` SSP-RK3 steps (main.f90:111-135) on the resident state.  dt of step k goes to
// dt_log[*first + k].  Without a time-dependent inflow nothing waits on the host; with one, the
// host reads each step's dt (while the first stage runs) to evaluate the boundary values.
int queue_steps(swe_handle* h, int nsteps, int* first) {
    const bool inflow = (h->bc_kind == SWE_BC_INFLOW);
    int rc;
    if (inflow && (rc = settle_time(h))) return rc;
    if ((rc = ensure_dt_log(h, nsteps))) return rc;
    const int base = h->pending;
    if (first) *first = base;
    const Mesh M = mesh_of(h);
    const unsigned g = grid_of(h->np);
    const bool lam = uses_lambda(h->prm);
    Node* A = h->S[0];
    Node* B = h->S[1];
    Node* Cc = h->S[2];
    for (int it = 0; it < nsteps; ++it) {
        timed_begin(h, "k_compute_dt");
        hipLaunchKernelGGL(k_compute_dt, dim3(g), dim3(kBlock), h->lds_bytes, h->stream, M, (const Node*)A,
                           lam ? h->d0 : (double*)nullptr, h->K, h->vmax, h->flags);
        timed_end(h, "k_compute_dt");
        hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(64), 0, h->stream, h->vmax, h->K.cfl, h->dtp, h->dt_log, base + it);
        hipEvent_t dt_ready = nullptr;
        if (inflow) {
            HIPCHK(hipMemcpyAsync(h->host_pinned, h->dtp, sizeof(double), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipEventCreateWithFlags(&dt_ready, hipEventDisableTiming));
            HIPCHK(hipEventRecord(dt_ready, h->stream));
        }
        // stage 1: U^n (A) -> B
        if ((rc = euler_stage(h, A, A, B, lam ? h->d0 : nullptr, 0))) return rc;
        if (inflow) {
            HIPCHK(hipEventSynchronize(dt_ready));
            HIPCHK(hipEventDestroy(dt_ready));
            const double dt = h->host_pinned[0];
            const double to = h->time;
            double t = to;
            launch_bdy(h, B, t + dt);
            t = to + dt;
            if ((rc = euler_stage(h, B, A, Cc, nullptr, 1))) return rc;
            launch_bdy(h, Cc, t + dt / 2);
            t = to + dt / 2;
            if ((rc = euler_stage(h, Cc, A, A, nullptr, 2))) return rc;
            launch_bdy(h, A, t + dt);
            h->time = to + dt;
        } else {
            launch_bdy(h, B, 0.0);
            if ((rc = euler_stage(h, B, A, Cc, nullptr, 1))) return rc;
            launch_bdy(h, Cc, 0.0);
            if ((rc = euler_stage(h, Cc, A, A, nullptr, 2))) return rc;
            launch_bdy(h, A, 0.0);
        }
    }
    if (!inflow) h->pending += nsteps;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SWE_EHIP, "kernel launch failed: %s", hipGetErrorString(e));
    return SWE_OK;
}

void free_handle(swe_handle* h) {
    if (!h) return;
    void* bufs[] = {h->len, h->sptr, h->perm, h->csr2slot, h->col, h->nvec, h->nrm, h->elist, h->eoff, h->dpos, h->c, h->m, h->z, h->S[0], h->S[1], h->S[2], h->scratch, h->ulo, h->un3,
                    h->d0, h->dL, h->muL, h->ework, h->alpha, h->res, h->Rm, h->nwork, h->vmax, h->dtp,
                    h->dt_log, h->flags, h->hl, h->uxl, h->uyl, h->hv, h->uxv, h->uyv};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (h->host_pinned) (void)hipHostFree(h->host_pinned);
    for (hipEvent_t e : h->tk.ev) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // namespace

// =============================================================================================
// C-ABI
// =============================================================================================
extern "C" {

const char* swe_last_error(void) { return g_err.c_str(); }

int swe_create(const swe_mesh_desc* d, const swe_params* prm, swe_handle** out) {
    g_err.clear();
    if (!d || !prm || !out) return fail(SWE_EINVAL, "null argument");
    *out = nullptr;
    if (d->np <= 0 || d->nnz <= 0 || !d->ia || !d->ja || !d->cx || !d->cy || !d->lumped || !d->bath)
        return fail(SWE_EINVAL, "mesh descriptor is incomplete");
    const int np = d->np, nnz = d->nnz;
    // validate the 1-based CSR on the host before anything reaches a kernel
    if (d->ia[0] != 1 || d->ia[np] - 1 != nnz) return fail(SWE_EINVAL, "ia must start at 1 and end at nnz+1");
    std::vector<int> rp((size_t)np + 1), cl((size_t)nnz);
    for (int i = 0; i <= np; ++i) {
        rp[(size_t)i] = d->ia[i] - 1;
        if (i > 0 && rp[(size_t)i] < rp[(size_t)i - 1]) return fail(SWE_EINVAL, "ia is not monotone at row %d", i);
    }
    for (int i = 0; i < np; ++i) {
        bool has_diag = false;
        for (int p = rp[(size_t)i]; p < rp[(size_t)i + 1]; ++p) {
            const int j = d->ja[p] - 1;
            if (j < 0 || j >= np) return fail(SWE_EINVAL, "ja(%d)=%d out of range", p + 1, d->ja[p]);
            cl[(size_t)p] = j;
            has_diag |= (j == i);
        }
        if (!has_diag) return fail(SWE_EINVAL, "row %d has no diagonal entry", i + 1);
    }
    for (int i = 0; i < np; ++i)
        if (!(d->lumped[i] > 0.0)) return fail(SWE_EINVAL, "lumped(%d) must be > 0", i + 1);
    auto check_list = [&](int n, const int* l, const char* name) -> int {
        if (n < 0 || (n > 0 && !l)) return fail(SWE_EINVAL, "%s list is malformed", name);
        for (int q = 0; q < n; ++q)
            if (l[q] < 1 || l[q] > np) return fail(SWE_EINVAL, "%s node %d out of range", name, l[q]);
        return SWE_OK;
    };
    int rc;
    if ((rc = check_list(d->n_h_dir, d->h_dir, "h_js_D"))) return rc;
    if ((rc = check_list(d->n_ux_dir, d->ux_dir, "ux_js_D"))) return rc;
    if ((rc = check_list(d->n_uy_dir, d->uy_dir, "uy_js_D"))) return rc;

    swe_handle* h = new swe_handle();
    h->np = np;
    h->nnz = nnz;
    h->prm = *prm;
    if ((rc = make_consts(*prm, h->K))) { delete h; return rc; }
    hipError_t e = hipGetDevice(&h->device);
    if (e != hipSuccess) { delete h; return fail(SWE_EHIP, "hipGetDevice: %s", hipGetErrorString(e)); }
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete h; return fail(SWE_EHIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    // Internal numbering: inside consecutive windows of kSortWindow rows of the caller's numbering,
    // rows are stably sorted by length, so the 64 rows of a slice (one wave) have equal length.
    // Only the processing order changes; each row keeps its entries in the caller's order.
    std::vector<int> perm((size_t)np), iperm((size_t)np);
    for (int w0 = 0; w0 < np; w0 += kSortWindow) {
        const int w1 = std::min(np, w0 + kSortWindow);
        for (int i = w0; i < w1; ++i) perm[(size_t)i] = i;
        std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int a, int b) {
            return (rp[(size_t)a + 1] - rp[(size_t)a]) < (rp[(size_t)b + 1] - rp[(size_t)b]);
        });
    }
    for (int r = 0; r < np; ++r) iperm[(size_t)perm[(size_t)r]] = r;
    const int np_pad = (np + 63) / 64 * 64;
    const int nsl = np_pad / 64;
    std::vector<int> len((size_t)np_pad, 0);
    std::vector<int64_t> sptr((size_t)nsl + 1, 0);
    for (int r = 0; r < np; ++r) len[(size_t)r] = rp[(size_t)perm[(size_t)r] + 1] - rp[(size_t)perm[(size_t)r]];
    for (int sl = 0; sl < nsl; ++sl) {
        int w = 0;
        for (int l = 0; l < 64; ++l) w = std::max(w, len[(size_t)sl * 64 + l]);
        sptr[(size_t)sl + 1] = sptr[(size_t)sl] + 64 * (int64_t)w;
    }
    const int64_t nslot = sptr[(size_t)nsl];
    int wmax = 0;
    for (int sl = 0; sl < nsl; ++sl) wmax = std::max(wmax, (int)((sptr[(size_t)sl + 1] - sptr[(size_t)sl]) / 64));
    if ((size_t)(2 + wmax) * 64 * sizeof(double) * (kBlock / 64) > 65536) {
        free_handle(h);
        return fail(SWE_EINVAL, "a row has %d entries; at most %d are supported", wmax,
                    (int)(65536 / (64 * sizeof(double) * (kBlock / 64))) - 2);
    }
    std::vector<int> col((size_t)nslot);
    std::vector<double2> cc((size_t)nslot, make_double2(0.0, 0.0));
    std::vector<double2> nv((size_t)nslot, make_double2(0.0, 0.0));
    std::vector<double> nr((size_t)nslot, 0.0);
    std::vector<int64_t> csr2slot((size_t)nnz);
    std::vector<double> m_int((size_t)np), z_int((size_t)np);
    for (int64_t q = 0; q < nslot; ++q) col[(size_t)q] = 0;
    std::vector<int> dpos((size_t)np, 0);
    for (int r = 0; r < np; ++r) {
        const int i = perm[(size_t)r];
        const int64_t base = sptr[(size_t)(r >> 6)] + (r & 63);
        for (int q = 0; q < len[(size_t)r]; ++q)
            if (cl[(size_t)rp[(size_t)i] + q] == i) dpos[(size_t)r] = q;
        for (int q = 0; q < len[(size_t)r]; ++q) {
            const int p = rp[(size_t)i] + q;
            const int64_t slot = base + 64 * (int64_t)q;
            col[(size_t)slot] = iperm[(size_t)cl[(size_t)p]];
            cc[(size_t)slot] = make_double2(d->cx[p], d->cy[p]);
            // compute_dij's normalisation (SOURCES/update.f90:306-310), same expressions and order
            const double norm = std::sqrt(0.0 + d->cx[p] * d->cx[p] + d->cy[p] * d->cy[p]);
            nr[(size_t)slot] = norm;
            nv[(size_t)slot] = make_double2(d->cx[p] / norm, d->cy[p] / norm);
            csr2slot[(size_t)p] = slot;
        }
        m_int[(size_t)r] = d->lumped[i];
        z_int[(size_t)r] = d->bath[i];
    }
    // compact per-slice lists of off-diagonal slots, in slot order (coalesced in the lambda phase)
    std::vector<int64_t> eoff((size_t)nsl + 1, 0);
    std::vector<uint16_t> elist;
    elist.reserve((size_t)(nnz - np));
    for (int sl = 0; sl < nsl; ++sl) {
        const int w = (int)((sptr[(size_t)sl + 1] - sptr[(size_t)sl]) / 64);
        for (int q = 0; q < w; ++q)
            for (int l = 0; l < 64; ++l) {
                const int r = sl * 64 + l;
                if (r < np && q < len[(size_t)r] && q != dpos[(size_t)r]) elist.push_back((uint16_t)(q * 64 + l));
            }
        eoff[(size_t)sl + 1] = (int64_t)elist.size();
    }
    h->np_pad = np_pad;
    h->nslot = nslot;
    h->wmax = wmax;
    h->lds_bytes = (size_t)(2 + wmax) * 64 * sizeof(double) * (kBlock / 64);
#define ALLOC(ptr, n) do { if ((rc = dalloc(&(ptr), (n)))) { free_handle(h); return rc; } } while (0)
    ALLOC(h->len, (size_t)np_pad);
    ALLOC(h->sptr, (size_t)nsl + 1);
    ALLOC(h->perm, (size_t)np);
    ALLOC(h->csr2slot, (size_t)nnz);
    ALLOC(h->col, (size_t)nslot);
    ALLOC(h->c, (size_t)nslot);
    ALLOC(h->elist, std::max<size_t>(elist.size(), 1));
    ALLOC(h->eoff, (size_t)nsl + 1);
    ALLOC(h->dpos, (size_t)np);
    ALLOC(h->nvec, (size_t)nslot);
    ALLOC(h->nrm, (size_t)nslot);
    ALLOC(h->m, (size_t)np);
    ALLOC(h->z, (size_t)np);
    for (int k = 0; k < 3; ++k) ALLOC(h->S[k], (size_t)np);
    ALLOC(h->scratch, (size_t)np);
    ALLOC(h->ulo, (size_t)np);
    ALLOC(h->un3, 3 * (size_t)np);
    ALLOC(h->d0, (size_t)nslot);
    ALLOC(h->dL, (size_t)nslot);
    ALLOC(h->muL, (size_t)nslot);
    ALLOC(h->ework, (size_t)std::max<int64_t>(nslot, nnz));
    ALLOC(h->alpha, (size_t)np);
    ALLOC(h->res, (size_t)np);
    ALLOC(h->Rm, (size_t)np);
    ALLOC(h->nwork, (size_t)np);
    ALLOC(h->vmax, 1);
    ALLOC(h->dtp, 1);
    ALLOC(h->flags, 1);
    ALLOC(h->hl, (size_t)std::max(d->n_h_dir, 1));
    ALLOC(h->uxl, (size_t)std::max(d->n_ux_dir, 1));
    ALLOC(h->uyl, (size_t)std::max(d->n_uy_dir, 1));
    ALLOC(h->hv, (size_t)std::max(d->n_h_dir, 1));
    ALLOC(h->uxv, (size_t)std::max(d->n_ux_dir, 1));
    ALLOC(h->uyv, (size_t)std::max(d->n_uy_dir, 1));
#undef ALLOC
    if (hipHostMalloc((void**)&h->host_pinned, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        free_handle(h);
        return fail(SWE_ENOMEM, "hipHostMalloc failed");
    }
    // Dirichlet lists in the internal numbering; values given later follow the caller's list order
    std::vector<int> lists[3];
    const int* src[3] = {d->h_dir, d->ux_dir, d->uy_dir};
    const int ns[3] = {d->n_h_dir, d->n_ux_dir, d->n_uy_dir};
    for (int k = 0; k < 3; ++k) {
        lists[k].resize((size_t)ns[k]);
        for (int q = 0; q < ns[k]; ++q) lists[k][(size_t)q] = iperm[(size_t)src[k][q] - 1];
    }
    h->nh = ns[0];
    h->nux = ns[1];
    h->nuy = ns[2];
    auto up = [&](void* dst, const void* s, size_t bytes) -> int {
        if (bytes == 0) return SWE_OK;
        hipError_t e2 = hipMemcpy(dst, s, bytes, hipMemcpyHostToDevice);
        if (e2 != hipSuccess) return fail(SWE_EHIP, "hipMemcpy: %s", hipGetErrorString(e2));
        return SWE_OK;
    };
    if ((rc = up(h->len, len.data(), sizeof(int) * len.size())) || (rc = up(h->sptr, sptr.data(), sizeof(int64_t) * sptr.size())) ||
        (rc = up(h->perm, perm.data(), sizeof(int) * perm.size())) ||
        (rc = up(h->csr2slot, csr2slot.data(), sizeof(int64_t) * csr2slot.size())) ||
        (rc = up(h->col, col.data(), sizeof(int) * col.size())) || (rc = up(h->c, cc.data(), sizeof(double2) * cc.size())) ||
        (rc = up(h->nvec, nv.data(), sizeof(double2) * nv.size())) ||
        (rc = up(h->elist, elist.data(), sizeof(uint16_t) * elist.size())) ||
        (rc = up(h->eoff, eoff.data(), sizeof(int64_t) * eoff.size())) || (rc = up(h->dpos, dpos.data(), sizeof(int) * dpos.size())) || (rc = up(h->nrm, nr.data(), sizeof(double) * nr.size())) ||
        (rc = up(h->m, m_int.data(), sizeof(double) * (size_t)np)) || (rc = up(h->z, z_int.data(), sizeof(double) * (size_t)np)) ||
        (rc = up(h->hl, lists[0].data(), sizeof(int) * lists[0].size())) ||
        (rc = up(h->uxl, lists[1].data(), sizeof(int) * lists[1].size())) ||
        (rc = up(h->uyl, lists[2].data(), sizeof(int) * lists[2].size()))) {
        free_handle(h);
        return rc;
    }
    e = hipMemset(h->vmax, 0, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(h->flags, 0, sizeof(int));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) { free_handle(h); return fail(SWE_EHIP, "init: %s", hipGetErrorString(e)); }
    *out = h;
    return SWE_OK;
}

int swe_destroy(swe_handle* h) {
    if (!h) return SWE_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    free_handle(h);
    return SWE_OK;
}

int swe_set_params(swe_handle* h, const swe_params* prm) {
    if (!h || !prm) return fail(SWE_EINVAL, "null argument");
    Consts K;
    int rc = make_consts(*prm, K);
    if (rc) return rc;
    h->prm = *prm;
    h->K = K;
    return SWE_OK;
}

int swe_set_bc(swe_handle* h, int kind, double amp, double period, double level) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if (kind < SWE_BC_NONE || kind > SWE_BC_VALUES) return fail(SWE_EINVAL, "unknown bc kind %d", kind);
    if (kind == SWE_BC_INFLOW && !(period != 0.0)) return fail(SWE_EINVAL, "inflow period must be non-zero");
    h->bc_kind = kind;
    h->bc_amp = amp;
    h->bc_period = period;
    h->bc_level = level;
    return SWE_OK;
}

int swe_set_bc_values(swe_handle* h, const double* hv, const double* uxv, const double* uyv) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if ((h->nh && !hv) || (h->nux && !uxv) || (h->nuy && !uyv)) return fail(SWE_EINVAL, "missing bc values");
    if (h->nh) HIPCHK(hipMemcpy(h->hv, hv, sizeof(double) * (size_t)h->nh, hipMemcpyHostToDevice));
    if (h->nux) HIPCHK(hipMemcpy(h->uxv, uxv, sizeof(double) * (size_t)h->nux, hipMemcpyHostToDevice));
    if (h->nuy) HIPCHK(hipMemcpy(h->uyv, uyv, sizeof(double) * (size_t)h->nuy, hipMemcpyHostToDevice));
    return SWE_OK;
}

int swe_set_state(swe_handle* h, const double* un) {
    if (!h || !un) return fail(SWE_EINVAL, "null argument");
    int rc = upload_state(h, un, h->S[0]);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    h->state_set = true;
    return SWE_OK;
}

int swe_get_state(swe_handle* h, double* un) {
    if (!h || !un) return fail(SWE_EINVAL, "null argument");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    return download_state(h, h->S[0], un);
}

int swe_compute_dij(swe_handle* h, const double* un, double* dij) {
    if (!h || !un || !dij) return fail(SWE_EINVAL, "null argument");
    int rc = upload_state(h, un, h->scratch);
    if (rc) return rc;
    hipLaunchKernelGGL(k_dij_full, dim3(grid_of(h->np)), dim3(kBlock), 0, h->stream, mesh_of(h),
                       (const Node*)h->scratch, h->dL, h->K, h->flags);
    hipLaunchKernelGGL(k_sell_to_csr, dim3((unsigned)((h->nnz + 255) / 256)), dim3(256), 0, h->stream,
                       (const double*)h->dL, (const int64_t*)h->csr2slot, h->ework, (int64_t)h->nnz);
    HIPCHK(hipMemcpyAsync(dij, h->ework, sizeof(double) * (size_t)h->nnz, hipMemcpyDeviceToHost, h->stream));
    return check_flags(h);
}

int swe_compute_dt(swe_handle* h, const double* un, double* dt) {
    if (!h || !dt) return fail(SWE_EINVAL, "null argument");
    const Node* U = h->S[0];
    if (un) {
        int rc = upload_state(h, un, h->scratch);
        if (rc) return rc;
        U = h->scratch;
    } else if (!h->state_set) {
        return fail(SWE_EINVAL, "no state has been set");
    }
    hipLaunchKernelGGL(k_compute_dt, dim3(grid_of(h->np)), dim3(kBlock), h->lds_bytes, h->stream, mesh_of(h), U,
                       (double*)nullptr, h->K, h->vmax, h->flags);
    hipLaunchKernelGGL(k_finalize_dt, dim3(1), dim3(64), 0, h->stream, h->vmax, h->K.cfl, h->dtp, (double*)nullptr, 0);
    HIPCHK(hipMemcpyAsync(h->host_pinned, h->dtp, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    int rc = check_flags(h);
    *dt = h->host_pinned[0];
    return rc;
}

int swe_euler(swe_handle* h, const double* un, double* unext, double dt) {
    if (!h || !un || !unext) return fail(SWE_EINVAL, "null argument");
    int rc = upload_state(h, un, h->scratch);
    if (rc) return rc;
    h->host_pinned[1] = dt;
    HIPCHK(hipMemcpyAsync(h->dtp, &h->host_pinned[1], sizeof(double), hipMemcpyHostToDevice, h->stream));
    Node* out = h->S[2];
    if ((rc = euler_stage(h, h->scratch, h->scratch, out, nullptr, 0))) return rc;
    if ((rc = download_state(h, out, unext))) return rc;
    return check_flags(h);
}

int swe_fct_positivity(swe_handle* h, const double* us, const double* f, double* lij) {
    if (!h || !us || !f || !lij) return fail(SWE_EINVAL, "null argument");
    const unsigned ge = (unsigned)((h->nnz + 255) / 256);
    HIPCHK(hipMemcpyAsync(h->un3, us, sizeof(double) * (size_t)h->np, hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(h->ework, f, sizeof(double) * (size_t)h->nnz, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_perm_nodes, dim3(grid_of(h->np)), dim3(kBlock), 0, h->stream, (const double*)h->un3,
                       (const int*)h->perm, h->nwork, h->np);
    hipLaunchKernelGGL(k_csr_to_sell, dim3(ge), dim3(256), 0, h->stream, (const double*)h->ework,
                       (const int64_t*)h->csr2slot, h->d0, (int64_t)h->nnz);
    const Mesh M = mesh_of(h);
    hipLaunchKernelGGL(k_fctpos_r, dim3(grid_of(h->np)), dim3(kBlock), 0, h->stream, M, (const double*)h->d0,
                       (const double*)h->nwork, h->Rm, h->K.htiny);
    hipLaunchKernelGGL(k_fctpos_l, dim3(grid_of(h->np)), dim3(kBlock), 0, h->stream, M, (const double*)h->d0,
                       (const double*)h->Rm, h->dL);
    hipLaunchKernelGGL(k_sell_to_csr, dim3(ge), dim3(256), 0, h->stream, (const double*)h->dL,
                       (const int64_t*)h->csr2slot, h->ework, (int64_t)h->nnz);
    HIPCHK(hipMemcpyAsync(lij, h->ework, sizeof(double) * (size_t)h->nnz, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return SWE_OK;
}

int swe_fct_positivity_csr(int np, int nnz, const int* ia, const int* ja, const double* us, const double* f,
                           double htiny, double* lij) {
    if (np <= 0 || nnz <= 0 || !ia || !ja || !us || !f || !lij) return fail(SWE_EINVAL, "null or empty argument");
    if (ia[0] != 1 || ia[np] - 1 != nnz) return fail(SWE_EINVAL, "ia must start at 1 and end at nnz+1");
    for (int i = 0; i < np; ++i)
        if (ia[i + 1] < ia[i]) return fail(SWE_EINVAL, "ia is not monotone at row %d", i + 1);
    for (int p = 0; p < nnz; ++p)
        if (ja[p] < 1 || ja[p] > np) return fail(SWE_EINVAL, "ja(%d)=%d out of range", p + 1, ja[p]);
    int *dia = nullptr, *dja = nullptr;
    double *dus = nullptr, *df = nullptr, *dR = nullptr, *dl = nullptr;
    int rc = SWE_OK;
    if ((rc = dalloc(&dia, (size_t)np + 1)) || (rc = dalloc(&dja, (size_t)nnz)) || (rc = dalloc(&dus, (size_t)np)) ||
        (rc = dalloc(&df, (size_t)nnz)) || (rc = dalloc(&dR, (size_t)np)) || (rc = dalloc(&dl, (size_t)nnz))) {
    } else {
        hipError_t e = hipMemcpy(dia, ia, sizeof(int) * ((size_t)np + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dja, ja, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dus, us, sizeof(double) * (size_t)np, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(df, f, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            const unsigned g = (unsigned)((np + 255) / 256);
            hipLaunchKernelGGL(k_fctpos_csr_r, dim3(g), dim3(256), 0, 0, np, (const int*)dia, (const double*)df,
                               (const double*)dus, dR, htiny);
            hipLaunchKernelGGL(k_fctpos_csr_l, dim3(g), dim3(256), 0, 0, np, (const int*)dia, (const int*)dja,
                               (const double*)df, (const double*)dR, dl);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpy(lij, dl, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(SWE_EHIP, "swe_fct_positivity_csr: %s", hipGetErrorString(e));
    }
    void* bufs[] = {dia, dja, dus, df, dR, dl};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    return rc;
}

int swe_set_time(swe_handle* h, double t) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    int rc = settle_time(h);
    if (rc) return rc;
    h->time = t;
    return SWE_OK;
}

int swe_get_time(swe_handle* h, double* t) {
    if (!h || !t) return fail(SWE_EINVAL, "null argument");
    int rc = settle_time(h);
    if (rc) return rc;
    *t = h->time;
    return SWE_OK;
}

int swe_run(swe_handle* h, int nsteps, double* time, double* dts) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    int rc;
    if (time && (rc = swe_set_time(h, *time))) return rc;
    int first = 0;
    if ((rc = queue_steps(h, nsteps, &first))) return rc;
    if (dts && nsteps > 0)
        HIPCHK(hipMemcpyAsync(dts, h->dt_log + first, sizeof(double) * (size_t)nsteps, hipMemcpyDeviceToHost, h->stream));
    const int frc = check_flags(h);
    if ((rc = settle_time(h))) return rc;
    if (time) *time = h->time;
    return frc;
}

int swe_run_async(swe_handle* h, int nsteps) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    return queue_steps(h, nsteps, nullptr);
}

int swe_sync(swe_handle* h) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    const int frc = check_flags(h);
    int rc = settle_time(h);
    return rc ? rc : frc;
}

void* swe_stream(swe_handle* h) { return h ? (void*)h->stream : nullptr; }

int swe_enable_kernel_timing(swe_handle* h, const char* kernel) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    h->timing = (kernel != nullptr && kernel[0] != 0);
    h->timed_name = h->timing ? kernel : "";
    h->tk.used = 0;
    return SWE_OK;
}

int swe_kernel_timing(swe_handle* h, const char* kernel, int* launches, double* ms) {
    if (!h || !kernel || !launches || !ms) return fail(SWE_EINVAL, "null argument");
    HIPCHK(hipStreamSynchronize(h->stream));
    *launches = 0;
    *ms = 0.0;
    if (h->timed_name != kernel) return SWE_OK;
    double sum = 0.0;
    for (int k = 0; k < h->tk.used; ++k) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, h->tk.ev[(size_t)2 * k], h->tk.ev[(size_t)2 * k + 1]));
        sum += t;
    }
    *launches = h->tk.used;
    *ms = sum;
    return SWE_OK;
}

}  // extern "C"
