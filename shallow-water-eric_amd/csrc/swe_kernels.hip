// swe_kernels.hip -- MI355X (gfx950) kernels of the explicit shallow-water step and their launch
// wrappers.
//
// One SSP-RK3 step (SOURCES/main.f90:71-91) runs as a short chain of row kernels over every
// instance of a pool at once.  A kernel gives each CSR row to one lane, walks the row's entries
// in their stored order and accumulates every sum sequentially, so each node's result is bitwise
// the reference's:
//
//   k_compute_dt    COMPUTE_DT + compute_dij  (main.f90:166-174, update.f90:291-326); keeps d_ij
//                   for the first stage and reduces max |d_ii|/m_i per instance
//   k_finalize_dt   dt = CFL/max per instance, the stage times of bdy, time = to + dt
//   k_alpha         alpha_limit, first loop   (update.f90:758-781)
//   k_stage_lo      first-order stage: compute_dij/compute_muij/MAX/alpha scaling, smb_2_roundoff
//                   (or smb_2, smb_1), sources, divide_by_lumped, update, RK combination
//                   (update.f90:180-223 and main.f90:83-89)
//   k_fct_lo        FCT stage, pass 1: low-order solution as above plus entropy_residual
//                   (update.f90:825-878) for the high-order viscosity
//   k_fct_rminus    FCT stage, pass 2: compute_fct_matrix + first loop of FCT_positivity
//                   (update.f90:721-748, fct.f90:15-30)
//   k_fct_apply     FCT stage, pass 3: l_ij (fct.f90:31-41), limited viscosities
//                   (update.f90:278-279), apply_viscosity (684-718), update, RK combination
//   k_ev_update     entropy-viscosity-only stage, pass 2 (update.f90:241-262)
//   k_bdy           bdy (main.f90:176-188)
//
// Every row kernel starts the same way: the lane's row, the wave's instance (one per 64-row
// slice, read once and kept in scalar registers), and an exit for free slices, instances that do
// not take this step and padding rows.
//
// Diagonal entries contribute to the reference's sums only through terms that multiply an exact
// zero (U_i*rat_ii - U_i*rat_ii); those products are skipped, which can change nothing but the
// sign of a zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "swe_engine.h"
#include "swe_kernels.h"

namespace swe {

namespace {

enum SmbKind : int { SMB_ROUNDOFF = 0, SMB_2 = 1, SMB_1 = 2, SMB_GALERKIN = 3 };

constexpr double kPi = 3.141592653589793;   // ACOS(-1.d0)

#define SWE_PRAGMA(x) _Pragma(#x)
#define SWE_UNROLL(n) SWE_PRAGMA(unroll n)
#ifndef SWE_UNROLL_LO
#define SWE_UNROLL_LO 1
#endif
#ifndef SWE_UNROLL_LO1
#define SWE_UNROLL_LO1 1
#endif
#ifndef SWE_UNROLL_RM
#define SWE_UNROLL_RM 1
#endif
#ifndef SWE_UNROLL_AP
#define SWE_UNROLL_AP 1
#endif
// minimum waves per SIMD requested from the register allocator (0: the compiler's choice).  The
// dt pass (VALU-bound, LDS-staged) runs best at 7 waves; k_fct_lo spills below 124 registers.
#ifndef SWE_WAVES_LO
#define SWE_WAVES_LO 0
#endif
#ifndef SWE_WAVES_DT
#define SWE_WAVES_DT 7
#endif
#if SWE_WAVES_LO > 0
#define SWE_LB_LO(BS) __launch_bounds__(BS, SWE_WAVES_LO)
#else
#define SWE_LB_LO(BS) __launch_bounds__(BS)
#endif
#ifndef SWE_WAVES_LO1
#define SWE_WAVES_LO1 5
#endif
#if SWE_WAVES_LO1 > 0
#define SWE_LB_LO1(BS) __launch_bounds__(BS, SWE_WAVES_LO1)
#else
#define SWE_LB_LO1(BS) __launch_bounds__(BS)
#endif
#if SWE_WAVES_DT > 0
#define SWE_LB_DT(BS) __launch_bounds__(BS, SWE_WAVES_DT)
#else
#define SWE_LB_DT(BS) __launch_bounds__(BS)
#endif

// Slice of this wave: the waves of a launch take the listed slices in order (remap: XCD-aware
// workgroup order), one slice each; -1 past the end of the list.  The value (and everything
// derived from it: the instance, its constants, dt) is wave-uniform and kept in scalar registers;
// the lane's row is slice*64 + lane.
template <int BS>
__device__ __forceinline__ int wave_slice(const Mesh& M, int remap) {
    const int b = remap ? xcd_block() : (int)blockIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(b * (BS / 64) + (int)(threadIdx.x >> 6));
    if (w >= M.nact) return -1;
    return __builtin_amdgcn_readfirstlane(M.act[w]);
}

// Instance owning a slice; -1 for a free slice.
__device__ __forceinline__ int slice_inst(const Mesh& M, int slice) {
    return __builtin_amdgcn_readfirstlane(M.sl[slice].inst);
}

__device__ __forceinline__ int lane_of() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ bool takes_steps(const InstDev& D, int flags) {
    return D.state == INST_RUNNING && flags == 0 && D.time < D.tfinal && D.steps < D.max_steps;
}

__device__ __forceinline__ int64_t row_base(const Mesh& M, int i) {
    return M.sl[i >> 6].sbase + (i & 63);
}

// Per-entry arrays streamed once per pass (c_ij, the first-order viscosities; with SWE_NT_LOADS >= 2
// also the dt pass's d_ij): their loads and stores carry the non-temporal hint, so they do not
// displace the gathered node data from the caches (measured: 14.86 -> 14.70 ms per 16.8M-node step).
// SWE_NT_NODE: the same hint on the low-order solution (1: discharge, 2: depth too), written by
// k_fct_lo and read back only at the row itself.
#ifndef SWE_NT_NODE
#define SWE_NT_NODE 2
#endif
#ifndef SWE_NT_LOADS
#define SWE_NT_LOADS 2
#endif
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld_entry(const double2* p) {
#if SWE_NT_LOADS
    const dvec2 v = __builtin_nontemporal_load((const dvec2*)p);
    return make_double2(v.x, v.y);
#else
    return *p;
#endif
}
__device__ __forceinline__ double ld_entry1(const double* p) {
#if SWE_NT_LOADS > 1
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}
__device__ __forceinline__ void st_entry1(double* p, double v) {
#if SWE_NT_LOADS > 1
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}
__device__ __forceinline__ void st_entry(double2* p, double2 v) {
#if SWE_NT_LOADS
    dvec2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, (dvec2*)p);
#else
    *p = v;
#endif
}

// Column of slot p of row i: a 16-bit offset from the row (half the bytes of the 32-bit column,
// which is read only where the offset does not fit).
__device__ __forceinline__ int col_at(const Mesh& M, int64_t p, int i) {
    const int d = M.dcol[p];
    return (d != kColFar) ? i + d : M.col[p];
}

__device__ __forceinline__ void raise_flag(int* flags, int inst, int f) {
    if (f) atomicOr(flags + inst, f);
}

// LDS of one wave for the compacted lambda phase: per-row one_over_h and SQRT(g*ABS(h)), then
// one double per slot of the slice (slot offset q*64 + lane).
__device__ __forceinline__ double* wave_lds(double* smem, int wmax) {
    return smem + (size_t)(threadIdx.x >> 6) * (size_t)(2 + wmax) * 64;
}

// RK combination of SOURCES/main.f90:83,89 (stage 0 keeps the Euler output)
__device__ __forceinline__ double rk_combine(int stage, double uo, double ui) {
    if (stage == 1) return (3.0 * uo + ui) / 4.0;
    if (stage == 2) return (uo + 2.0 * ui) / 3.0;
    return ui;
}

// Source terms of one node, dispatched on type_test (update.f90:614-623): friction (425-452),
// my_friction (454-469), coriolis (471-486).  v1, v2 = velocity at U.  src is uniform per wave.
// Launches whose instances have no source term use the HAS_SRC = false kernels, which leave it out.
template <bool HAS_SRC>
__device__ __forceinline__ void add_sources(int src, const Node& I, double mi, double v1, double v2, double dt,
                                            const Consts& K, double& r2, double& r3) {
    if constexpr (!HAS_SRC) return;
    if (src == SRC_MANNING) {
        const double hs = pow43(fmaxF(I.h, K.htiny));
        const double vel = K.fric * sqrt_rn(v1 * v1 + v2 * v2);
        const double hloc = 1.0 * vel * dt;
        const double den = hs + fmaxF(hs, hloc);
        r2 = r2 - mi * I.qx * vel * 2.0 / den;
        r3 = r3 - mi * I.qy * vel * 2.0 / den;
    } else if (src == SRC_LINEAR) {
        r2 = r2 - mi * 0.2 * I.qx;
        r3 = r3 - mi * 0.2 * I.qy;
    } else if (src == SRC_CORIOLIS) {
        const double f = 2.0;
        r2 = r2 + mi * f * I.qy;
        r3 = r3 - mi * f * I.qx;
    }
}

// compute_dij of one off-diagonal entry from c_ij itself: the norm and the unit vector are
// evaluated with compute_dij's expressions (SOURCES/update.f90:306-310; correctly rounded sqrt and
// division, so the bits are the reference's), which costs a square root and two divisions per entry
// instead of storing 24 bytes of normal and norm per entry (measured faster in every pass).
__device__ __forceinline__ double dij_from_c(const double2 c, const Node& I, const Node& J, double ovh_i,
                                             double ovh_j, double sql, double sqr, const Consts& K, int& err) {
    const double norm = sqrt_normal(0.0 + c.x * c.x + c.y * c.y);
    return dij_of(make_double2(c.x / norm, c.y / norm), norm, I, J, ovh_i, ovh_j, sql, sqr, K, err);
}

// dlam is compute_dij's value for this entry (from the dt pass or computed inline).
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

// alpha_limit, first loop (update.f90:758-781) for row i: alpha_i = (|sum(h_j-h_i)| /
// sum|h_j-h_i|)**8, 1 if h_i <= htiny or the denominator does not exceed the numerator.  All index
// loads, then all gathers, then the sums in entry order.  (Evaluating it inside the dt pass for
// stage 1 was measured slower than this separate sweep: the gathers after the LDS phase stall.)
__device__ __forceinline__ double alpha_of(const Mesh& M, const Node* __restrict__ U, int i, double hi, double htiny) {
    double a = 1.0;
    if (!(hi <= htiny)) {
        const int64_t pb = row_base(M, i);
        const int len = M.len[i];
        constexpr int kUnroll = 12;
        int js[kUnroll];
        double hs[kUnroll];
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) js[q] = (q < len) ? (int)M.dcol[pb + 64 * q] : 0;
#pragma unroll
        for (int q = 0; q < kUnroll; ++q) js[q] = (js[q] != kColFar) ? i + js[q] : M.col[pb + 64 * q];
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
            const double hj = U[col_at(M, pb + 64 * (int64_t)q, i)].h;
            num = num + hj - hi;
            denom = denom + fabs(hj - hi);
        }
        if (denom <= fabs(num)) a = 1.0;
        else a = fabs(num) / (denom);
        a = a * a;
        a = a * a;
        a = a * a;
    }
    return a;
}

// =============================================================================================
// COMPUTE_DT: d_ij for every off-diagonal entry, row sums, max |d_ii|/m_i per instance.
// Phase 0 (one lane per row) puts the row's one_over_h and SQRT(g*ABS(h)) in LDS and an exact 0.0
// in the diagonal slot (what the reference's row sum adds there); phase 1 (entry-parallel over
// the slice's compact off-diagonal list) evaluates d_ij = |c_ij| lambda_ij into LDS (and d0);
// phase 2 sums each row in entry order.  all = 1 (host-array calls) covers every admitted
// instance; otherwise only the instances that take the step.
// =============================================================================================
// The dt pass on one slice per wave (slice -1: a wave past the end of the list, which still takes
// part in the workgroup's barriers); every wave of the workgroup calls it together.
template <int BS>
__device__ __forceinline__ void dt_slice(const Mesh& M, const InstDev* __restrict__ T, const int* __restrict__ flags,
                                         const Node* __restrict__ U, double* __restrict__ d0,
                                         unsigned long long* __restrict__ vmax, int* __restrict__ flags_out, int all,
                                         int slice, double* lw) {
    const int lane = lane_of();
    bool wave_live = false;
    int inst = -1;
    if (slice >= 0) {
        inst = slice_inst(M, slice);
        if (inst >= 0) wave_live = all ? (T[inst].state != INST_FREE) : takes_steps(T[inst], flags[inst]);
    }
    const int i = slice * 64 + lane;
    const bool valid = wave_live && M.len[i] > 0;
    int err = 0;
    Consts K{};
    if (wave_live) K = T[inst].K;
    Node I{};
    if (valid) {
        I = U[i];
        lw[lane] = ovh_of(I.h, K.htiny);
        lw[64 + lane] = sqrt_rn(K.g * fabs(I.h));
        lw[128 + 64 * M.dpos[i] + lane] = 0.0;
    }
    __syncthreads();
    if (wave_live) {
        const SliceInfo S = M.sl[slice];
        const int64_t e1 = S.ebase + S.ecount;
        for (int64_t e = S.ebase + lane; e < e1; e += 64) {
            const int off = M.elist[e];
            const int r = off & 63;
            const int64_t slot = S.sbase + off;
            const Node Ir = U[slice * 64 + r];
            const Node J = U[col_at(M, slot, slice * 64 + r)];
            const double d = dij_from_c(ld_entry(M.c + slot), Ir, J, lw[r], ovh_of(J.h, K.htiny), lw[64 + r],
                                        sqrt_rn(K.g * fabs(J.h)), K, err);
            lw[128 + off] = d;
            if (d0) st_entry1(d0 + slot, d);
        }
    }
    __syncthreads();
    double v = 0.0;
    if (valid) {
        // -SUM(dij%aa(ia(i):ia(i+1)-1)) with the diagonal slot holding 0.0, in entry order
        double s = 0.0;
        const int len = M.len[i];
        for (int q = 0; q < len; ++q) s = s + lw[128 + 64 * q + lane];
        v = fabs(-s) / M.m[i];
    }
    if (err) atomicOr(flags_out + inst, err);
    // wave max, then per block one atomic per run of waves of the same instance (v >= 0 orders
    // like its bit pattern)
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(b, off, 64);
        b = (o > b) ? o : b;
    }
    __shared__ unsigned long long red[BS / 64];
    __shared__ int rinst[BS / 64];
    if (lane == 0) {
        red[threadIdx.x >> 6] = b;
        rinst[threadIdx.x >> 6] = wave_live ? inst : -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int cur = -1;
        unsigned long long mx = 0ull;
        for (int w = 0; w < BS / 64; ++w) {
            if (rinst[w] != cur) {
                if (cur >= 0) atomicMax(vmax + cur, mx);
                cur = rinst[w];
                mx = red[w];
            } else {
                mx = (red[w] > mx) ? red[w] : mx;
            }
        }
        if (cur >= 0) atomicMax(vmax + cur, mx);
    }
}

template <int BS>
__global__ void SWE_LB_DT(BS)
k_compute_dt(Mesh M, const InstDev* __restrict__ T, const int* __restrict__ flags, const Node* __restrict__ U,
             double* __restrict__ d0, unsigned long long* __restrict__ vmax, int* __restrict__ flags_out, int all,
             int remap) {
    extern __shared__ double smem[];
    dt_slice<BS>(M, T, flags, U, d0, vmax, flags_out, all, wave_slice<BS>(M, remap), wave_lds(smem, M.wmax));
}

// Per-instance step bookkeeping after the dt pass (SOURCES/main.f90:61-91):
// an instance takes the step while time < tfinal (and its step limit and error flags allow);
// dt = CFL*1/MAXVAL(ABS(d_ii)/m_i); the boundary values of the three stages are evaluated at
// t+dt, (t+dt)+dt/2 and (t+dt/2)+dt; time = to + dt.  advance = 0 only computes dt.
__device__ __forceinline__ void finalize_one(InstDev* __restrict__ T, unsigned long long* __restrict__ vmax,
                                             const int* __restrict__ flags, int k, int advance) {
    InstDev& D = T[k];
    if (D.state == INST_FREE) return;
    const double mx = __longlong_as_double((long long)vmax[k]);
    vmax[k] = 0ull;
    if (!advance) {
        D.dt = D.K.cfl * 1.0 / mx;
        return;
    }
    if (D.state != INST_RUNNING) {
        D.go = 0;
        return;
    }
    if (!takes_steps(D, flags[k])) {
        D.go = 0;
        D.state = flags[k] ? INST_FAILED : INST_DONE;
        return;
    }
    // test 9 takes the driver's fixed dt and skips COMPUTE_DT (main.f90:66-71)
    const double dt = (D.K.dt_fixed > 0.0) ? D.K.dt_fixed : D.K.cfl * 1.0 / mx;
    const double to = D.time;
    D.dt = dt;
    if (D.bc_kind == SWE_BC_INFLOW) {
        double t = to;
        const double ts0 = t + dt;
        t = to + dt;
        const double ts1 = t + dt / 2;
        t = to + dt / 2;
        const double ts2 = t + dt;
        D.bc_val[0] = D.bc_amp * sin_rn(2.0 * kPi * ts0 / D.bc_period);
        D.bc_val[1] = D.bc_amp * sin_rn(2.0 * kPi * ts1 / D.bc_period);
        D.bc_val[2] = D.bc_amp * sin_rn(2.0 * kPi * ts2 / D.bc_period);
    }
    if (D.dt_log) D.dt_log[D.steps % D.dt_cap] = dt;
    D.time = to + dt;
    D.steps = D.steps + 1;
    D.go = 1;
}

__global__ void k_finalize_dt(InstDev* __restrict__ T, unsigned long long* __restrict__ vmax,
                              const int* __restrict__ flags, int ninst, int advance) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < ninst) finalize_one(T, vmax, flags, k, advance);
}

// compute_dij with the full dij%aa (diagonal = -row sum), for swe_compute_dij.
template <int BS>
__global__ void __launch_bounds__(BS)
k_dij_full(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, double* __restrict__ dij,
           int* __restrict__ flags_out) {
    const int slice = wave_slice<BS>(M, 0);
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || T[inst].state == INST_FREE) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    const Consts K = T[inst].K;
    int err = 0;
    const Node I = U[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double sql = sqrt_rn(K.g * fabs(I.h));
    double s = 0.0;
    int64_t pd = -1;
    const int64_t pb = row_base(M, i);
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        if (j == i) { pd = p; dij[p] = 0.0; continue; }
        const Node J = U[j];
        const double d = dij_from_c(M.c[p], I, J, ovh_i, ovh_of(J.h, K.htiny), sql, sqrt_rn(K.g * fabs(J.h)), K, err);
        dij[p] = d;
        s = s + d;
    }
    if (pd >= 0) dij[pd] = -s;
    raise_flag(flags_out, inst, err);
}

// =============================================================================================
// alpha_limit, first loop: alpha_i = (|sum(h_j-h_i)| / sum|h_j-h_i|)**8, 1 if dry.
// =============================================================================================
__device__ __forceinline__ void alpha_slice(const Mesh& M, const InstDev* __restrict__ T, const Node* __restrict__ U,
                                            double* __restrict__ alpha, int slice) {
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || !T[inst].go) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    alpha[i] = alpha_of(M, U, i, U[i].h, T[inst].K.htiny);
}

template <int BS>
__global__ void __launch_bounds__(BS)
k_alpha(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, double* __restrict__ alpha, int remap) {
    alpha_slice(M, T, U, alpha, wave_slice<BS>(M, remap));
}

// =============================================================================================
// First-order stage ('viscous', or 'galerkin'): one pass per stage.
// =============================================================================================
template <int BS, int SMB, bool ALPHA, bool HAS_SRC>
__global__ void __launch_bounds__(BS)
k_stage_lo(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, const Node* __restrict__ Uo,
           Node* __restrict__ out, const double* __restrict__ d0, const double* __restrict__ alpha, int stage,
           int* __restrict__ flags_out, int remap) {
    const int slice = wave_slice<BS>(M, remap);
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || !T[inst].go) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    const Consts K = T[inst].K;
    const double dt = T[inst].dt;
    int err = 0;
    const Node I = U[i];
    const double sql = sqrt_rn(K.g * fabs(I.h));
    const double mi = M.m[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double v1i = I.qx * ovh_i, v2i = I.qy * ovh_i;
    const double ai = ALPHA ? alpha[i] : 1.0;
    const double thr = (SMB == SMB_ROUNDOFF || SMB == SMB_1) ? K.htiny : 0.0;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0, fix = 0.0;
    const int64_t pb = row_base(M, i);
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        const Node J = U[j];
        const double2 c = ld_entry(M.c + p);
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
                                   d0 ? d0[p] : dij_from_c(c, I, J, ovh_i, ovh_j, sql, sqrt_rn(K.g * fabs(J.h)), K, err),
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
    add_sources<HAS_SRC>(K.src, I, mi, v1i, v2i, dt, K, r2, r3);
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
    raise_flag(flags_out, inst, err);
}

// =============================================================================================
// First-order stage of the SGN model (type_test 13: 5 components h, qx, qy, h*eta, h*w; 'viscous'
// or 'galerkin').  The same sums as k_stage_lo over every component (smb_2_roundoff, smb_2, smb_1 of
// update.f90:366-423, 554-682 with syst_size = 5), with the model's flux (condlim.f90:52-75: the
// second velocity is v*0, so only x-fluxes remain; lambda and mu keep the full discharge), then
// mSGN_RHS (update.f90:488-551) as a second pass over the row, the update of every component and
// the RK combination.  One kernel per stage; its workgroup size and slice order (dispatch or
// XCD-aware) are the tuned ones of the first-order stage, whose place it takes in the step.
// =============================================================================================
// pTilde (update.f90:515-518) at node n: paper_constant = lambdaSGN*g/(3*SQRT(lumped)),
// x = (h*eta)/h**2; the literal integer powers are products from the left, as flang expands them.
__device__ __forceinline__ double sgn_ptilde(double h, double he, double m, const Consts& K) {
    const double pc = K.lambda_sgn * K.g / (3.0 * sqrt_rn(m));
    const double x = he / (h * h);
    return pc * (h * h * h) * (2.0 + 4.0 * (x * x * x) - 6.0 * (x * x * x * x));
}

template <int BS, int SMB, bool ALPHA>
__global__ void __launch_bounds__(BS)
k_stage_sgn(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, const NodeX* __restrict__ UX,
            const Node* __restrict__ Uo, const NodeX* __restrict__ UoX, Node* __restrict__ out,
            NodeX* __restrict__ outX, const double* __restrict__ d0, const double* __restrict__ alpha, int stage,
            int* __restrict__ flags_out, int remap) {
    const int slice = wave_slice<BS>(M, remap);
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || !T[inst].go) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    const Consts K = T[inst].K;
    const double dt = T[inst].dt;
    int err = 0;
    const Node I = U[i];
    const NodeX IX = UX[i];
    const double sql = sqrt_rn(K.g * fabs(I.h));
    const double mi = M.m[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double v1i = I.qx * ovh_i, v2i = I.qy * ovh_i;
    const double v2fi = I.qy * ovh_i * 0.0;   // velocity(2,i) of the model's flux
    const double ui[5] = {I.h, I.qx, I.qy, IX.he, IX.hw};
    const double ai = ALPHA ? alpha[i] : 1.0;
    const double thr = (SMB == SMB_ROUNDOFF || SMB == SMB_1) ? K.htiny : 0.0;
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double fix = 0.0;
    const int64_t pb = row_base(M, i);
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        const Node J = U[j];
        const NodeX JX = UX[j];
        const double uj[5] = {J.h, J.qx, J.qy, JX.he, JX.hw};
        const double2 c = M.c[p];
        const double ovh_j = ovh_of(J.h, K.htiny);
        const double v1j = J.qx * ovh_j, v2j = J.qy * ovh_j;
        const double v2fj = J.qy * ovh_j * 0.0;
        const double zmax = fmaxF(I.z, J.z);
        const double Hij = fmaxF(0.0, I.h + I.z - zmax);
        const double Hji = fmaxF(0.0, J.h + J.z - zmax);
        const double ratij = (I.h <= thr) ? 0.0 : Hij / I.h;
        const double ratji = (J.h <= thr) ? 0.0 : Hji / J.h;
        double d = 0.0, mu = 0.0;
        const bool off = (j != i);
        if (SMB != SMB_GALERKIN && off)
            first_order_dmu<ALPHA>(c, v1i, v2i, v1j, v2j,
                                   d0 ? d0[p] : dij_from_c(c, I, J, ovh_i, ovh_j, sql, sqrt_rn(K.g * fabs(J.h)), K, err),
                                   ai, alpha, j, d, mu);
        if constexpr (SMB == SMB_1) {
            for (int k = 0; k < 5; ++k) {
                double xx = 0.0;
                xx = xx - c.x * (v1j * uj[k] * ratji + v1i * ui[k] * ratij);
                xx = xx - c.y * (v2fj * uj[k] * ratji + v2fi * ui[k] * ratij);
                r[k] = r[k] + xx;
                if (off) r[k] = r[k] + d * (uj[k] * ratji - ui[k] * ratij);
            }
            if (off) {
                const double t = 0.5 * K.g * (Hji * Hji - Hij * Hij);
                r[1] = r[1] - t * c.x;
                r[2] = r[2] - t * c.y;
            }
        } else {
            int k0 = 0;
            if constexpr (SMB == SMB_ROUNDOFF) {
                double xm = 0.0;
                xm = xm - c.x * (v1j);
                xm = xm - c.y * (v2fj);
                if (!off) fix = fix + xm;
                else {
                    fix = fix - mu + (d - mu) * (-ratij);
                    r[0] = r[0] + (xm + mu) * J.h + fmaxF((d - mu) * Hji, 0.0);
                }
                k0 = 1;
            }
            for (int k = k0; k < 5; ++k) {
                double xx = 0.0;
                xx = xx - c.x * (v1j * uj[k]);
                xx = xx - c.y * (v2fj * uj[k]);
                r[k] = r[k] + xx;
                if (SMB != SMB_GALERKIN && off)
                    r[k] = r[k] + d * (uj[k] * ratji - ui[k] * ratij) + mu * (uj[k] * (1.0 - ratji) - ui[k] * (1.0 - ratij));
            }
            const double topo = K.g * I.h * (J.h + J.z - I.h - I.z);
            r[1] = r[1] - topo * c.x;
            r[2] = r[2] - topo * c.y;
        }
    }
    // mSGN_RHS: the momenta lose pTilde(j)*c_ij over the row in CSR order, then the h*eta and h*w sources
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        const double2 c = M.c[p];
        const double pt = sgn_ptilde(U[j].h, UX[j].he, M.m[j], K);
        r[1] = r[1] - pt * c.x;
        r[2] = r[2] - pt * c.y;
    }
    {
        const double pc = K.lambda_sgn * K.g / (3.0 * sqrt_rn(mi));
        const double x = IX.he / (I.h * I.h);
        const double psi = 12.0 * (x - 1.0);
        const double t = IX.he / I.h;
        const double s = 3.0 * pc * (t * t) * psi;
        r[3] = r[3] + mi * IX.hw;
        r[4] = r[4] - mi * s;
    }
    for (int k = 0; k < 5; ++k) r[k] = r[k] / mi;
    double o[5];
    if constexpr (SMB == SMB_1) o[0] = I.h + dt * r[0];
    else if constexpr (SMB == SMB_ROUNDOFF) o[0] = I.h * (1.0 + dt * fix / mi) + dt * r[0];
    else o[0] = fmaxF(I.h + dt * r[0], 0.0);
    for (int k = 1; k < 5; ++k) o[k] = ui[k] + dt * r[k];
    if (stage != 0) {
        const Node P = Uo[i];
        const NodeX PX = UoX[i];
        const double po[5] = {P.h, P.qx, P.qy, PX.he, PX.hw};
        for (int k = 0; k < 5; ++k) o[k] = rk_combine(stage, po[k], o[k]);
    }
    Node w;
    w.h = o[0];
    w.qx = o[1];
    w.qy = o[2];
    w.z = I.z;
    out[i] = w;
    outX[i] = NodeX{o[3], o[4]};
    raise_flag(flags_out, inst, err);
}

// =============================================================================================
// FCT / entropy-viscosity stage, pass 1: first-order viscosities (kept in dL, muL), the
// low-order solution (smb_2_roundoff, update) and the entropy residual res_i.
// =============================================================================================
// Where the row loop takes d_ij = |c_ij| lambda_ij (compute_dij at the stage's state) from:
enum LamSrc : int {
    LAM_INLINE = 0,   // evaluated per entry inside the row loop (lane per row)
    LAM_D0 = 1,       // the dt pass's d_ij at U (stage 1), read from HBM
    LAM_LDS = 2,      // evaluated entry-parallel into the wave's LDS before the row loop
};

template <int BS, bool ALPHA, bool WRITE_LO, bool HAS_SRC, int LAM>
__device__ __forceinline__ void
fct_lo_body(const Mesh& M, const InstDev* __restrict__ T, const Node* __restrict__ U, double* __restrict__ ulo_h,
            double2* __restrict__ ulo_q,
            double2* __restrict__ dmu, double* __restrict__ res, const double* __restrict__ d0,
            const double* __restrict__ lam, const double* __restrict__ alpha, int roundoff,
            int* __restrict__ flags_out, int slice) {
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || !T[inst].go) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    const Consts K = T[inst].K;
    const double dt = T[inst].dt;
    int err = 0;
    const Node I = U[i];
    const double sql = sqrt_rn(K.g * fabs(I.h));
    const double mi = M.m[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double v1i = I.qx * ovh_i, v2i = I.qy * ovh_i;
    const double ai = ALPHA ? alpha[i] : 1.0;
    const double thr_lo = roundoff ? K.htiny : 0.0;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0, fix = 0.0;   // smb_2_roundoff (or smb_2)
    double e1 = 0.0, e2 = 0.0, e3 = 0.0;              // smb_2 of entropy_residual
    double emax = 0.0, emin = 0.0;
    bool split = false;   // the entropy residual's momentum sums have left the low-order ones
    const int64_t pb = row_base(M, i);
    constexpr int kUnroll = (LAM == LAM_D0) ? SWE_UNROLL_LO1 : SWE_UNROLL_LO;
#pragma unroll kUnroll
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        const Node J = U[j];
        const double2 c = ld_entry(M.c + p);
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
            // LAM_D0 / LAM_LDS: the row loop carries no lambda code and needs fewer registers
            // (5 waves per SIMD instead of 4)
            double dlam;
            if constexpr (LAM == LAM_D0) dlam = ld_entry1(d0 + p);
            else if constexpr (LAM == LAM_LDS) dlam = lam[64 * q + lane_of()];
            else dlam = dij_from_c(c, I, J, ovh_i, ovh_j, sql, sqrt_rn(K.g * fabs(J.h)), K, err);
            first_order_dmu<ALPHA>(c, v1i, v2i, v1j, v2j, dlam, ai, alpha, j, d, mu);
            st_entry(dmu + p, make_double2(d, mu));
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
        // The momentum sums of the low-order update (thresholds htiny) and of the entropy residual
        // (thresholds 0) take the same terms in the same order while every entry's ratios agree, so
        // they are the same numbers: e2, e3 are copied from r2, r3 at the first entry whose ratios
        // differ (or at the end) instead of being summed alongside.
        if (WRITE_LO && !split && !(ratjir == ratji0 && ratijr == ratij0)) {
            split = true;
            e2 = r2;
            e3 = r3;
        }
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
        if (off) e1 = e1 + d * (J.h * ratji0 - I.h * ratij0) + mu * (J.h * (1.0 - ratji0) - I.h * (1.0 - ratij0));
        if (!WRITE_LO || split) {
            e2 = e2 + xx2;
            e3 = e3 + xx3;
            if (off) {
                e2 = e2 + d * (J.qx * ratji0 - I.qx * ratij0) + mu * (J.qx * (1.0 - ratji0) - I.qx * (1.0 - ratij0));
                e3 = e3 + d * (J.qy * ratji0 - I.qy * ratij0) + mu * (J.qy * (1.0 - ratji0) - I.qy * (1.0 - ratij0));
            }
            e2 = e2 - topo * c.x;
            e3 = e3 - topo * c.y;
        }
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
    if (WRITE_LO && !split) {
        e2 = r2;
        e3 = r3;
    }
    add_sources<HAS_SRC>(K.src, I, mi, v1i, v2i, dt, K, e2, e3);
    if (WRITE_LO) {
        add_sources<HAS_SRC>(K.src, I, mi, v1i, v2i, dt, K, r2, r3);
        r1 = r1 / mi;
        r2 = r2 / mi;
        r3 = r3 / mi;
        // the low-order solution, depth and discharge apart: the R- pass reads the depth only
        double hL;
        if (roundoff) hL = I.h * (1.0 + dt * fix / mi) + dt * r1;
        else hL = fmaxF(I.h + dt * r1, 0.0);
#if SWE_NT_NODE > 1
        st_entry1(ulo_h + i, hL);
#else
        ulo_h[i] = hL;
#endif
#if SWE_NT_NODE > 0
        st_entry(ulo_q + i, make_double2(I.qx + dt * r2, I.qy + dt * r3));
#else
        ulo_q[i] = make_double2(I.qx + dt * r2, I.qy + dt * r3);
#endif
        if (K.check_h && hL < 0.0) err |= FLAG_NEG_H;
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
    raise_flag(flags_out, inst, err);
}

// Stages 2 and 3 (lambda inline) and stage 1 (d_ij of the dt pass; without the lambda code the
// kernel fits SWE_WAVES_LO1 waves per SIMD).
template <int BS, bool ALPHA, bool WRITE_LO, bool HAS_SRC>
__global__ void SWE_LB_LO(BS)
k_fct_lo(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, double* __restrict__ ulo_h,
         double2* __restrict__ ulo_q, double2* __restrict__ dmu, double* __restrict__ res,
         const double* __restrict__ alpha, int roundoff, int* __restrict__ flags_out, int remap) {
    fct_lo_body<BS, ALPHA, WRITE_LO, HAS_SRC, LAM_INLINE>(M, T, U, ulo_h, ulo_q, dmu, res, nullptr, nullptr, alpha,
                                                          roundoff, flags_out, wave_slice<BS>(M, remap));
}

template <int BS, bool ALPHA, bool WRITE_LO, bool HAS_SRC>
__global__ void SWE_LB_LO1(BS)
k_fct_lo1(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, double* __restrict__ ulo_h,
          double2* __restrict__ ulo_q, double2* __restrict__ dmu, double* __restrict__ res,
          const double* __restrict__ d0, const double* __restrict__ alpha, int roundoff, int* __restrict__ flags_out,
          int remap) {
    fct_lo_body<BS, ALPHA, WRITE_LO, HAS_SRC, LAM_D0>(M, T, U, ulo_h, ulo_q, dmu, res, d0, nullptr, alpha, roundoff,
                                                      flags_out, wave_slice<BS>(M, remap));
}

// Stages 2 and 3 with lambda entry-parallel (schedule kind 1): the wave first evaluates
// d_ij = |c_ij| lambda_ij of its slice's off-diagonal entries, one entry per lane over the slice's
// compact entry list as the dt pass does, into its LDS; then the row loop runs as in k_fct_lo1 with
// d_ij read from LDS.  Every d_ij is the same function of the same operands as in the inline form,
// so the results are the same bits; lambda no longer idles the lanes of diagonal and padding slots.
template <int BS, bool ALPHA, bool WRITE_LO, bool HAS_SRC>
__global__ void SWE_LB_LO1(BS)
k_fct_lo_lds(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, double* __restrict__ ulo_h,
             double2* __restrict__ ulo_q, double2* __restrict__ dmu, double* __restrict__ res,
             const double* __restrict__ alpha, int roundoff, int* __restrict__ flags_out, int remap) {
    extern __shared__ double smem[];
    double* lw = wave_lds(smem, M.wmax);
    const int slice = wave_slice<BS>(M, remap);
    const int lane = lane_of();
    int inst = -1;
    bool live = false;
    if (slice >= 0) {
        inst = slice_inst(M, slice);
        live = inst >= 0 && T[inst].go;
    }
    if (live) {
        const int i = slice * 64 + lane;
        if (M.len[i] > 0) {
            const double h = U[i].h;
            lw[lane] = ovh_of(h, T[inst].K.htiny);
            lw[64 + lane] = sqrt_rn(T[inst].K.g * fabs(h));
        }
    }
    __syncthreads();
    if (live) {
        const Consts K = T[inst].K;
        const SliceInfo S = M.sl[slice];
        const int64_t e1 = S.ebase + S.ecount;
        int err = 0;
        for (int64_t e = S.ebase + lane; e < e1; e += 64) {
            const int off = M.elist[e];
            const int r = off & 63;
            const int64_t slot = S.sbase + off;
            const Node Ir = U[slice * 64 + r];
            const Node J = U[col_at(M, slot, slice * 64 + r)];
            lw[128 + off] = dij_from_c(ld_entry(M.c + slot), Ir, J, lw[r], ovh_of(J.h, K.htiny), lw[64 + r],
                                       sqrt_rn(K.g * fabs(J.h)), K, err);
        }
        raise_flag(flags_out, inst, err);
    }
    __syncthreads();
    fct_lo_body<BS, ALPHA, WRITE_LO, HAS_SRC, LAM_LDS>(M, T, U, ulo_h, ulo_q, dmu, res, nullptr, lw + 128, alpha,
                                                       roundoff, flags_out, slice);
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
__device__ __forceinline__ void rminus_slice(const Mesh& M, const InstDev* __restrict__ T, const Node* __restrict__ U,
                                             const double* __restrict__ ulo_h, const double2* __restrict__ dmu,
                                             const double* __restrict__ res, double* __restrict__ Rm, int slice) {
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || !T[inst].go) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    const double dt = T[inst].dt;
    const double htiny = T[inst].K.htiny;
    const Node I = U[i];
    const double resi = res[i];
    const double scaled = (M.m[i] / dt) * ulo_h[i];
    const double Q = 0.0 - scaled;
    double P = -1.0e-15;
    const int64_t pb = row_base(M, i);
    SWE_UNROLL(SWE_UNROLL_RM)
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        if (j == i) continue;    // f_ii is a zero, never < 0
        const Node J = U[j];
        double dH, muH, rij_, rji_;
        const double2 lo = ld_entry(dmu + p);
        const double f = fct_flux(I, J, lo.x, lo.y, fmaxF(resi, res[j]), dH, muH, rij_, rji_);
        if (f < 0.0) P = P + f;
    }
    double R = fminF(Q / P, 1.0);
    if (scaled <= htiny) R = 0.0;
    Rm[i] = R;
}

template <int BS>
__global__ void __launch_bounds__(BS)
k_fct_rminus(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, const double* __restrict__ ulo_h,
             const double2* __restrict__ dmu, const double* __restrict__ res,
             double* __restrict__ Rm, int remap) {
    rminus_slice(M, T, U, ulo_h, dmu, res, Rm, wave_slice<BS>(M, remap));
}

// =============================================================================================
// FCT stage, pass 3: l_ij, limited viscosities, apply_viscosity, update, RK combination.
// =============================================================================================
__device__ __forceinline__ void apply_slice(const Mesh& M, const InstDev* __restrict__ T, const Node* __restrict__ U,
                                            const Node* __restrict__ Uo, const double* __restrict__ ulo_h,
                                            const double2* __restrict__ ulo_q, Node* __restrict__ out,
                                            const double2* __restrict__ dmu, const double* __restrict__ res,
                                            const double* __restrict__ Rm, int stage, int* __restrict__ flags_out,
                                            int slice) {
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || !T[inst].go) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    const double dt = T[inst].dt;
    const int check_h = T[inst].K.check_h;
    const Node I = U[i];
    const double resi = res[i];
    const double Ri = Rm[i];
    double x1 = 0.0, x2 = 0.0, x3 = 0.0;
    const int64_t pb = row_base(M, i);
    SWE_UNROLL(SWE_UNROLL_AP)
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        if (j == i) continue;    // contributes d*(0) + mu*(0)
        const Node J = U[j];
        const double2 lo = ld_entry(dmu + p);
        const double dLp = lo.x, muLp = lo.y;
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
#if SWE_NT_NODE > 1
    const double hL = ld_entry1(ulo_h + i);   // low-order solution of k_fct_lo
#else
    const double hL = ulo_h[i];       // low-order solution of k_fct_lo
#endif
#if SWE_NT_NODE > 0
    const double2 qL = ld_entry(ulo_q + i);
#else
    const double2 qL = ulo_q[i];
#endif
    const double h = hL + dt * (x1 / mi);
    const double qx = qL.x + dt * (x2 / mi);
    const double qy = qL.y + dt * (x3 / mi);
    const int err = (check_h && h < 0.0) ? FLAG_NEG_H : 0;
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
    raise_flag(flags_out, inst, err);
}

template <int BS>
__global__ void __launch_bounds__(BS)
k_fct_apply(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, const Node* __restrict__ Uo,
            const double* __restrict__ ulo_h, const double2* __restrict__ ulo_q, Node* __restrict__ out,
            const double2* __restrict__ dmu, const double* __restrict__ res, const double* __restrict__ Rm, int stage,
            int* __restrict__ flags_out, int remap) {
    apply_slice(M, T, U, Uo, ulo_h, ulo_q, out, dmu, res, Rm, stage, flags_out, wave_slice<BS>(M, remap));
}

// =============================================================================================
// Entropy-viscosity-only stage, pass 2 (update.f90:241-262).
// =============================================================================================
template <int BS, bool HAS_SRC>
__global__ void __launch_bounds__(BS)
k_ev_update(Mesh M, const InstDev* __restrict__ T, const Node* __restrict__ U, const Node* __restrict__ Uo,
            Node* __restrict__ out, const double2* __restrict__ dmu,
            const double* __restrict__ res, int roundoff, int stage, int* __restrict__ flags_out, int remap) {
    const int slice = wave_slice<BS>(M, remap);
    if (slice < 0) return;
    const int inst = slice_inst(M, slice);
    if (inst < 0 || !T[inst].go) return;
    const int i = slice * 64 + lane_of();
    const int len = M.len[i];
    if (len == 0) return;
    const Consts K = T[inst].K;
    const double dt = T[inst].dt;
    const Node I = U[i];
    const double mi = M.m[i];
    const double ovh_i = ovh_of(I.h, K.htiny);
    const double v1i = I.qx * ovh_i, v2i = I.qy * ovh_i;
    const double resi = res[i];
    const double thr = roundoff ? K.htiny : 0.0;
    double r1 = 0.0, r2 = 0.0, r3 = 0.0, fix = 0.0;
    const int64_t pb = row_base(M, i);
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        const int j = col_at(M, p, i);
        const Node J = U[j];
        const double2 c = ld_entry(M.c + p);
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
            const double2 lo = ld_entry(dmu + p);
            d = fminF(lo.x, 1.0 * rij);
            mu = fminF(lo.y, 1.0 * rij);
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
    add_sources<HAS_SRC>(K.src, I, mi, v1i, v2i, dt, K, r2, r3);
    r1 = r1 / mi;
    r2 = r2 / mi;
    r3 = r3 / mi;
    double h;
    if (roundoff) h = I.h * (1.0 + dt * fix / mi) + dt * r1;
    else h = fmaxF(I.h + dt * r1, 0.0);
    const double qx = I.qx + dt * r2;
    const double qy = I.qy + dt * r3;
    const int err = (K.check_h && h < 0.0) ? FLAG_NEG_H : 0;
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
    raise_flag(flags_out, inst, err);
}

// =============================================================================================
// bdy: Dirichlet overwrites after a stage, for every instance that takes the step.
// =============================================================================================
__device__ __forceinline__ void bdy_one(Node* __restrict__ u, const BcEntry* __restrict__ E, int q,
                                        const double* __restrict__ vals, int64_t stride,
                                        const InstDev* __restrict__ T, int stage) {
    const BcEntry e = E[q];
    if (e.inst < 0) return;
    const InstDev& D = T[e.inst];
    if (!D.go) return;
    const int kind = D.bc_kind;
    if (kind == SWE_BC_NONE) return;
    Node& x = u[e.node];
    // SWE_BC_VALUES: one value per entry; SWE_BC_STAGES: one per entry and stage
    const bool given = kind == SWE_BC_VALUES || kind == SWE_BC_STAGES;
    const double v = given ? vals[(kind == SWE_BC_STAGES ? stage * stride : 0) + e.vidx] : 0.0;
    if (e.comp == 0) {
        if (kind == SWE_BC_INFLOW) x.h = fmaxF(D.bc_level - x.z + D.bc_val[stage], 0.0);
        else if (given) x.h = v;
    } else if (e.comp == 1) {
        x.qx = v;
    } else {
        x.qy = v;
    }
}

__global__ void k_bdy(Node* __restrict__ u, const BcEntry* __restrict__ E, int n, const double* __restrict__ vals,
                      int64_t stride, const InstDev* __restrict__ T, int stage) {
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (q < n) bdy_one(u, E, q, vals, stride, T, stage);
}

// =============================================================================================
// Persistent FCT step: every pass of `nsteps` steps in one cooperative launch, the passes
// separated by grid-wide barriers instead of kernel boundaries (the same passes, the same per-slice
// code and the same order as the launch sequence of one step, so the same bits).  Meant for
// meshes too small to fill the GPU, where a dependent kernel launch costs more than the pass.
// The barrier wait is bounded: past kBarrierSpins polls a workgroup raises the abort word and every
// workgroup leaves, so the launch always ends; k_step_check then fails every stepping instance with
// FLAG_STEP.  (The cooperative launch places every workgroup at once, so the bound is a guard.)
// =============================================================================================
struct StepArgs {
    Mesh M;
    InstDev* T;
    int* flags;
    Node *S0, *S1, *S2;
    double* d0;
    double* alpha;
    double* ulo_h;
    double2* ulo_q;
    double2* dmu;
    double* res;
    double* Rm;
    unsigned long long* vmax;
    const BcEntry* bc;
    const double* bcv;
    int64_t bcv_stride;
    int nbc;
    int ninst;
    int roundoff;
    unsigned* bar;      // bar[0]: barrier arrivals since the launch; bar[1]: abort word
    long spin_limit;    // polls a barrier waits at most (kBarrierSpins; SWE_BARRIER_SPINS overrides)
};

constexpr long kBarrierSpins = 1L << 21;

// All threads of every workgroup call it together.  Each thread first waits until its own memory
// operations have completed (they are in the L2 of its XCD); then one thread per workgroup
// releases at device scope (the L2 written back, so the other XCDs see it), arrives, waits for
// the other workgroups and acquires at device scope (the CU's L1 and the L2 invalidated), before
// the workgroup goes on.
__device__ __forceinline__ bool grid_barrier(const StepArgs& A, unsigned& epoch) {
    __shared__ int ok;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    epoch += 1;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_s_waitcnt(0);   // the write-back is complete before the arrival counts
        const unsigned target = epoch * gridDim.x;
        // relaxed: the fences before and after order the workgroup's data
        __hip_atomic_fetch_add(A.bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int good = 1;
        long spins = 0;
        while (__hip_atomic_load(A.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(A.bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                good = 0;
                break;
            }
            if (++spins > A.spin_limit) {
                __hip_atomic_fetch_or(A.bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __builtin_amdgcn_s_waitcnt(0);
        ok = good;
    }
    __syncthreads();
    return ok != 0;
}

#ifndef SWE_WAVES_STEP
#define SWE_WAVES_STEP 2
#endif
template <bool ALPHA, bool HAS_SRC>
__global__ void __launch_bounds__(256, SWE_WAVES_STEP)
k_step_fct(StepArgs A, int nsteps) {
    constexpr int BS = 256, WPB = BS / 64;
    extern __shared__ double smem[];
    double* lw = wave_lds(smem, A.M.wmax);
    const Mesh& M = A.M;
    const int wave = (int)(threadIdx.x >> 6);
    const int nw = (int)gridDim.x * WPB;
    const int gw = (int)blockIdx.x * WPB + wave;
    const int gt = (int)(blockIdx.x * BS + threadIdx.x), nt = (int)gridDim.x * BS;
    unsigned epoch = 0;
    for (int step = 0; step < nsteps; ++step) {
        // COMPUTE_DT: whole workgroups take rounds of WPB slices (the pass synchronises a workgroup)
        for (int base = (int)blockIdx.x * WPB; base < M.nact; base += nw) {
            const int w = base + wave;
            const int slice = (w < M.nact) ? __builtin_amdgcn_readfirstlane(M.act[w]) : -1;
            dt_slice<BS>(M, A.T, A.flags, A.S0, A.d0, A.vmax, A.flags, 0, slice, lw);
        }
        if (!grid_barrier(A, epoch)) return;
        for (int k = gt; k < A.ninst; k += nt) finalize_one(A.T, A.vmax, A.flags, k, 1);
        if (!grid_barrier(A, epoch)) return;
        for (int stage = 0; stage < 3; ++stage) {
            const Node* U = (stage == 0) ? A.S0 : (stage == 1) ? A.S1 : A.S2;
            Node* out = (stage == 0) ? A.S1 : (stage == 1) ? A.S2 : A.S0;
            if (ALPHA) {
                for (int w = gw; w < M.nact; w += nw)
                    alpha_slice(M, A.T, U, A.alpha, __builtin_amdgcn_readfirstlane(M.act[w]));
                if (!grid_barrier(A, epoch)) return;
            }
            for (int w = gw; w < M.nact; w += nw) {
                const int slice = __builtin_amdgcn_readfirstlane(M.act[w]);
                if (stage == 0)
                    fct_lo_body<BS, ALPHA, true, HAS_SRC, LAM_D0>(M, A.T, U, A.ulo_h, A.ulo_q, A.dmu, A.res, A.d0,
                                                                  nullptr, A.alpha, A.roundoff, A.flags, slice);
                else
                    fct_lo_body<BS, ALPHA, true, HAS_SRC, LAM_INLINE>(M, A.T, U, A.ulo_h, A.ulo_q, A.dmu, A.res,
                                                                      nullptr, nullptr, A.alpha, A.roundoff, A.flags,
                                                                      slice);
            }
            if (!grid_barrier(A, epoch)) return;
            for (int w = gw; w < M.nact; w += nw)
                rminus_slice(M, A.T, U, A.ulo_h, A.dmu, A.res, A.Rm, __builtin_amdgcn_readfirstlane(M.act[w]));
            if (!grid_barrier(A, epoch)) return;
            for (int w = gw; w < M.nact; w += nw)
                apply_slice(M, A.T, U, A.S0, A.ulo_h, A.ulo_q, out, A.dmu, A.res, A.Rm, stage, A.flags,
                            __builtin_amdgcn_readfirstlane(M.act[w]));
            if (!grid_barrier(A, epoch)) return;
            for (int q = gt; q < A.nbc; q += nt) bdy_one(out, A.bc, q, A.bcv, A.bcv_stride, A.T, stage);
            if (!grid_barrier(A, epoch)) return;
        }
    }
}

// After a persistent launch: if a barrier gave up, every running instance fails (its step may have
// stopped anywhere).
__global__ void k_step_check(const unsigned* __restrict__ bar, const InstDev* __restrict__ T, int* __restrict__ flags,
                             int ninst) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k < ninst && bar[1] != 0u && T[k].state == INST_RUNNING) atomicOr(flags + k, FLAG_STEP);
}

// =============================================================================================
// Layout conversion and table edits
// =============================================================================================
// un(3,np) in the caller's numbering <-> internal Node records of one instance; perm maps its
// internal rows to the caller's nodes.
__global__ void k_pack(const double* __restrict__ un3, const double* __restrict__ z, const int* __restrict__ perm,
                       Node* __restrict__ u, int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np) return;
    const int k = perm[r];
    Node o;
    o.h = un3[3 * (size_t)k];
    o.qx = un3[3 * (size_t)k + 1];
    o.qy = un3[3 * (size_t)k + 2];
    o.z = z[r];
    u[r] = o;
}

// un(5,np) of the SGN model <-> Node records and their (h*eta, h*w) beside them
__global__ void k_pack5(const double* __restrict__ un5, const double* __restrict__ z, const int* __restrict__ perm,
                        Node* __restrict__ u, NodeX* __restrict__ x, int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np) return;
    const size_t k = (size_t)perm[r];
    Node o;
    o.h = un5[5 * k];
    o.qx = un5[5 * k + 1];
    o.qy = un5[5 * k + 2];
    o.z = z[r];
    u[r] = o;
    x[r] = NodeX{un5[5 * k + 3], un5[5 * k + 4]};
}

__global__ void k_unpack5(const Node* __restrict__ u, const NodeX* __restrict__ x, const int* __restrict__ perm,
                          double* __restrict__ un5, int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np) return;
    const size_t k = (size_t)perm[r];
    const Node o = u[r];
    const NodeX e = x[r];
    un5[5 * k] = o.h;
    un5[5 * k + 1] = o.qx;
    un5[5 * k + 2] = o.qy;
    un5[5 * k + 3] = e.he;
    un5[5 * k + 4] = e.hw;
}

__global__ void k_unpack(const Node* __restrict__ u, const int* __restrict__ perm, double* __restrict__ un3, int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= np) return;
    const int k = perm[r];
    const Node o = u[r];
    un3[3 * (size_t)k] = o.h;
    un3[3 * (size_t)k + 1] = o.qx;
    un3[3 * (size_t)k + 2] = o.qy;
}

__global__ void k_perm_nodes(const double* __restrict__ src, const int* __restrict__ perm, double* __restrict__ dst,
                             int np) {
    const int r = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < np) dst[r] = src[perm[r]];
}

__global__ void k_csr_to_sell(const double* __restrict__ src, const int64_t* __restrict__ map, double* __restrict__ dst,
                              int64_t nnz) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) dst[map[p]] = src[p];
}

__global__ void k_sell_to_csr(const double* __restrict__ src, const int64_t* __restrict__ map, double* __restrict__ dst,
                              int64_t nnz) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nnz) dst[p] = src[map[p]];
}

__global__ void k_fill_slices(SliceInfo* __restrict__ sl, int64_t s0, int64_t n, int inst) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) sl[s0 + k].inst = inst;
}

// Small table edits ride in the kernel arguments, so queuing one never waits for the stream.
__global__ void k_poke(uint32_t* __restrict__ dst, PokeWords w, int nwords) {
    const int k = threadIdx.x;
    if (k < nwords) dst[k] = w.w[k];
}

__global__ void k_fill_bc(BcEntry* __restrict__ bc, int64_t b0, int64_t n) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) bc[b0 + k].inst = -1;
}

// FCT_positivity on the pool's sliced layout (fct.f90:15-41): R- per row, then l_ij.
__global__ void k_fctpos_r(Mesh M, const double* __restrict__ f, const double* __restrict__ us, double htiny,
                           double* __restrict__ Rm) {
    const int slice = wave_slice<256>(M, 0);
    if (slice < 0) return;
    const int i = slice * 64 + lane_of();
    if (M.len[i] == 0) return;
    const double Q = 0.0 - us[i];
    double P = -1.0e-15;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    for (int q = 0; q < len; ++q) {
        const double fij = f[pb + 64 * (int64_t)q];
        if (fij < 0.0) P = P + fij;
    }
    double R = fminF(Q / P, 1.0);
    if (us[i] <= htiny) R = 0.0;
    Rm[i] = R;
}

__global__ void k_fctpos_l(Mesh M, const double* __restrict__ f, const double* __restrict__ Rm, double* __restrict__ lij) {
    const int slice = wave_slice<256>(M, 0);
    if (slice < 0) return;
    const int i = slice * 64 + lane_of();
    if (M.len[i] == 0) return;
    const int64_t pb = row_base(M, i);
    const int len = M.len[i];
    const double Ri = Rm[i];
    for (int q = 0; q < len; ++q) {
        const int64_t p = pb + 64 * (int64_t)q;
        lij[p] = (f[p] >= 0.0) ? Rm[col_at(M, p, i)] : Ri;
    }
}

// The same on a plain 1-based CSR matrix (swe_fct_positivity_csr and its variant without
// allocations).  Row ranges are clipped to [0, nnz) and columns outside 1..np give l_ij = NaN, so
// a malformed pattern that the host could not check (device arrays) cannot make a kernel read
// out of bounds.
__global__ void k_fctpos_csr_r(int np, int nnz, const int* __restrict__ ia, const double* __restrict__ f,
                               const double* __restrict__ us, double* __restrict__ Rm, double htiny) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const double Q = 0.0 - us[i];
    double P = -1.0e-15;
    const int p0 = max(ia[i] - 1, 0), p1 = min(ia[i + 1] - 1, nnz);
    for (int p = p0; p < p1; ++p) {
        const double fij = f[p];
        if (fij < 0.0) P = P + fij;
    }
    double R = fminF(Q / P, 1.0);
    if (us[i] <= htiny) R = 0.0;
    Rm[i] = R;
}

__global__ void k_fctpos_csr_l(int np, int nnz, const int* __restrict__ ia, const int* __restrict__ ja,
                               const double* __restrict__ f, const double* __restrict__ Rm, double* __restrict__ lij) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const int p0 = max(ia[i] - 1, 0), p1 = min(ia[i + 1] - 1, nnz);
    for (int p = p0; p < p1; ++p) {
        const int j = ja[p] - 1;
        lij[p] = (j < 0 || j >= np) ? __builtin_nan("") : (f[p] >= 0.0) ? Rm[j] : Rm[i];
    }
}

// FCT_generic (SOURCES/SAFE/fct.f90:3-43), the two-sided limiter on a plain 1-based CSR matrix:
// Q+- = dg*(max/min - u), P+- = +-1e-15 + sums of the positive / negative fluxes in entry order,
// R+- = MIN(Q/P, 1); then u_i += (SUM_p l_ij f_ij)/dg_i with l_ij = MIN(R+_i, R-_j) for f_ij >= 0,
// MIN(R-_i, R+_j) otherwise.
__global__ void k_fctgen_csr_r(int np, const int* __restrict__ ia, const double* __restrict__ f,
                               const double* __restrict__ u, const double* __restrict__ maxn,
                               const double* __restrict__ minn, const double* __restrict__ dg,
                               double2* __restrict__ R) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const double Qp = dg[i] * (maxn[i] - u[i]);
    const double Qm = dg[i] * (minn[i] - u[i]);
    double Pp = 1.0e-15, Pm = -1.0e-15;
    for (int p = ia[i] - 1; p < ia[i + 1] - 1; ++p) {
        const double fij = f[p];
        if (fij >= 0.0) Pp = Pp + fij;
        else Pm = Pm + fij;
    }
    R[i] = make_double2(fminF(Qp / Pp, 1.0), fminF(Qm / Pm, 1.0));
}

__global__ void k_fctgen_csr_apply(int np, const int* __restrict__ ia, const int* __restrict__ ja,
                                   const double* __restrict__ f, const double2* __restrict__ R,
                                   const double* __restrict__ dg, double* __restrict__ u) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= np) return;
    const double2 Ri = R[i];
    double x = 0.0;
    for (int p = ia[i] - 1; p < ia[i + 1] - 1; ++p) {
        const double fij = f[p];
        const double2 Rj = R[ja[p] - 1];
        const double lij = (fij >= 0.0) ? fminF(Ri.x, Rj.y) : fminF(Ri.y, Rj.x);
        x = x + lij * fij;
    }
    u[i] = u[i] + x / dg[i];
}

inline unsigned blocks_of(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Dispatch a kernel template over the supported workgroup sizes.
#define SWE_BS_SWITCH(bs, ...)                    \
    switch (bs) {                                 \
        case 128: {                               \
            constexpr int BS = 128;               \
            __VA_ARGS__;                          \
        } break;                                  \
        case 512: {                               \
            constexpr int BS = 512;               \
            __VA_ARGS__;                          \
        } break;                                  \
        default: {                                \
            constexpr int BS = 256;               \
            __VA_ARGS__;                          \
        } break;                                  \
    }

void timed_begin(Pool& P, int kid) {
    if (!P.timing || P.timed_name != kKernelNames[kid]) return;
    TimedKernel& T = P.tk;
    if ((size_t)(2 * T.used + 2) > T.ev.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            T.ev.push_back(e);
        }
    }
    (void)hipEventRecord(T.ev[(size_t)2 * T.used], P.stream);
}

void timed_end(Pool& P, int kid) {
    if (!P.timing || P.timed_name != kKernelNames[kid]) return;
    TimedKernel& T = P.tk;
    (void)hipEventRecord(T.ev[(size_t)2 * T.used + 1], P.stream);
    T.used++;
}

template <int SMB, bool ALPHA>
void launch_stage_lo(Pool& P, const Node* U, const Node* Uo, Node* out, const double* d0, int stage) {
    const Mesh M = mesh_view(P);
    const LaunchCfg& L = P.tune.k[KID_STAGE_LO];
    SWE_BS_SWITCH(L.block, {
        const dim3 g(blocks_of((int64_t)M.nact * 64, BS));
        if (P.any_src)
            hipLaunchKernelGGL((k_stage_lo<BS, SMB, ALPHA, true>), g, dim3(BS), 0, P.stream, M, (const InstDev*)P.tab,
                               U, Uo, out, d0, (const double*)P.alpha, stage, P.flags, L.remap);
        else
            hipLaunchKernelGGL((k_stage_lo<BS, SMB, ALPHA, false>), g, dim3(BS), 0, P.stream, M, (const InstDev*)P.tab,
                               U, Uo, out, d0, (const double*)P.alpha, stage, P.flags, L.remap);
    })
}

template <bool ALPHA, bool WRITE_LO>
void launch_fct_lo(Pool& P, const Node* U, const double* d0) {
    const Mesh M = mesh_view(P);
    const LaunchCfg& L = P.tune.k[KID_FCT_LO];
    const int ro = P.mode.roundoff;
    auto go = [&](auto src, auto has_d0) {
        constexpr bool S = decltype(src)::value, D = decltype(has_d0)::value;
        if (!D && L.kind == 1) {
            // entry-parallel lambda: (2 + wmax)*64 doubles of LDS per wave, within 64 KiB per workgroup
            int block = L.block;
            while (block > 128 && (size_t)(2 + M.wmax) * 64 * sizeof(double) * (block / 64) > 65536) block /= 2;
            SWE_BS_SWITCH(block, {
                const size_t lds = (size_t)(2 + M.wmax) * 64 * sizeof(double) * (BS / 64);
                hipLaunchKernelGGL((k_fct_lo_lds<BS, ALPHA, WRITE_LO, S>), dim3(blocks_of((int64_t)M.nact * 64, BS)),
                                   dim3(BS), lds, P.stream, M, (const InstDev*)P.tab, U, P.ulo_h, P.ulo_q, P.dmu,
                                   P.res, (const double*)P.alpha, ro, P.flags, L.remap);
            })
        } else if constexpr (D) {
            SWE_BS_SWITCH(L.block, hipLaunchKernelGGL((k_fct_lo1<BS, ALPHA, WRITE_LO, S>),
                                                      dim3(blocks_of((int64_t)M.nact * 64, BS)), dim3(BS), 0, P.stream,
                                                      M, (const InstDev*)P.tab, U, P.ulo_h, P.ulo_q, P.dmu, P.res, d0,
                                                      (const double*)P.alpha, ro, P.flags, L.remap))
        } else {
            SWE_BS_SWITCH(L.block, hipLaunchKernelGGL((k_fct_lo<BS, ALPHA, WRITE_LO, S>),
                                                      dim3(blocks_of((int64_t)M.nact * 64, BS)), dim3(BS), 0, P.stream,
                                                      M, (const InstDev*)P.tab, U, P.ulo_h, P.ulo_q, P.dmu, P.res,
                                                      (const double*)P.alpha, ro, P.flags, L.remap))
        }
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    if (P.any_src) {
        if (d0) go(T_{}, T_{});
        else go(T_{}, F_{});
    } else {
        if (d0) go(F_{}, T_{});
        else go(F_{}, F_{});
    }
}

}  // namespace

const char* const kKernelNames[KID_COUNT] = {"k_compute_dt", "k_alpha", "k_stage_lo", "k_fct_lo",
                                             "k_fct_rminus", "k_fct_apply", "k_ev_update"};

int kinds_of(int kid) { return (kid == KID_FCT_LO) ? 2 : 1; }

bool block_supported(int kid, int block) {
    if (kid == KID_STAGE_LO || kid == KID_EV) return block == 256;
    return block == 128 || block == 256 || block == 512;
}

Mesh mesh_view(const Pool& P) {
    return Mesh{P.sl, P.len, P.col, P.dcol, P.c, P.m, P.elist, P.dpos, P.act, P.nact, P.wmax};
}

void launch_compute_dt(Pool& P, const Node* U, double* d0, int all) {
    const Mesh M = mesh_view(P);
    if (M.nact == 0) return;
    const LaunchCfg& L = P.tune.k[KID_DT];
    // the dynamic LDS of a workgroup stays within 64 KiB: fall back to smaller workgroups
    int block = L.block;
    while (block > 128 && (size_t)(2 + M.wmax) * 64 * sizeof(double) * (block / 64) > 65536) block /= 2;
    timed_begin(P, KID_DT);
    SWE_BS_SWITCH(block, {
        const size_t lds = (size_t)(2 + M.wmax) * 64 * sizeof(double) * (BS / 64);
        hipLaunchKernelGGL((k_compute_dt<BS>), dim3(blocks_of((int64_t)M.nact * 64, BS)), dim3(BS), lds, P.stream, M,
                           (const InstDev*)P.tab, (const int*)P.flags, U, d0, P.vmax, P.flags, all, L.remap);
    })
    timed_end(P, KID_DT);
}

void launch_finalize_dt(Pool& P, int advance) {
    const int n = P.ninst();
    if (n == 0) return;
    hipLaunchKernelGGL(k_finalize_dt, dim3(blocks_of(n, 64)), dim3(64), 0, P.stream, P.tab, P.vmax,
                       (const int*)P.flags, n, advance);
}

// One Euler stage from U to out (out != U) for every instance that takes the step.  d0 (may be
// null) holds compute_dij at U.  stage: 0 keeps the Euler output, 1 and 2 apply the RK
// combinations with Uo.
int launch_euler_stage(Pool& P, const Node* U, const Node* Uo, Node* out, const double* d0, int stage) {
    const Mode& md = P.mode;
    const Mesh M = mesh_view(P);
    if (M.nact == 0) return SWE_OK;
    const InstDev* T = P.tab;
    const bool alpha = md.alpha != 0 && md.viscosity_type != SWE_GALERKIN;
    if (alpha) {
        const LaunchCfg& L = P.tune.k[KID_ALPHA];
        timed_begin(P, KID_ALPHA);
        SWE_BS_SWITCH(L.block, hipLaunchKernelGGL((k_alpha<BS>), dim3(blocks_of((int64_t)M.nact * 64, BS)), dim3(BS), 0, P.stream,
                                                  M, T, U, P.alpha, L.remap))
        timed_end(P, KID_ALPHA);
    }
    if (md.sgn) {
        // the SGN model (first-order and galerkin schemes only)
        const NodeX* UX = ext_of(P, U);
        const NodeX* UoX = ext_of(P, Uo);
        NodeX* outX = ext_of(P, out);
        if (!UX || !UoX || !outX) return fail(SWE_EINVAL, "SGN stage on a buffer without h*eta, h*w");
        const LaunchCfg& L = P.tune.k[KID_STAGE_LO];
        const int smb = (md.viscosity_type == SWE_GALERKIN) ? SMB_GALERKIN
                        : (md.viscous_type == 1)             ? SMB_1
                        : (md.roundoff ? SMB_ROUNDOFF : SMB_2);
        timed_begin(P, KID_STAGE_LO);
#define SWE_SGN_LAUNCH(S, A)                                                                                       \
    hipLaunchKernelGGL((k_stage_sgn<BS, S, A>), dim3(blocks_of((int64_t)M.nact * 64, BS)), dim3(BS), 0, P.stream, M, \
                       T, U, UX, Uo, UoX, out, outX, (S == SMB_GALERKIN) ? nullptr : d0, (const double*)P.alpha,     \
                       stage, P.flags, L.remap)
        SWE_BS_SWITCH(L.block, {
            if (smb == SMB_GALERKIN) SWE_SGN_LAUNCH(SMB_GALERKIN, false);
            else if (smb == SMB_1) {
                if (alpha) SWE_SGN_LAUNCH(SMB_1, true);
                else SWE_SGN_LAUNCH(SMB_1, false);
            } else if (smb == SMB_ROUNDOFF) {
                if (alpha) SWE_SGN_LAUNCH(SMB_ROUNDOFF, true);
                else SWE_SGN_LAUNCH(SMB_ROUNDOFF, false);
            } else {
                if (alpha) SWE_SGN_LAUNCH(SMB_2, true);
                else SWE_SGN_LAUNCH(SMB_2, false);
            }
        })
#undef SWE_SGN_LAUNCH
        timed_end(P, KID_STAGE_LO);
    } else if (md.viscosity_type == SWE_GALERKIN) {
        launch_stage_lo<SMB_GALERKIN, false>(P, U, Uo, out, nullptr, stage);
    } else if (md.viscosity_type == SWE_VISCOUS) {
        const int smb = (md.viscous_type == 1) ? SMB_1 : (md.roundoff ? SMB_ROUNDOFF : SMB_2);
        timed_begin(P, KID_STAGE_LO);
        if (smb == SMB_1) {
            if (alpha) launch_stage_lo<SMB_1, true>(P, U, Uo, out, d0, stage);
            else launch_stage_lo<SMB_1, false>(P, U, Uo, out, d0, stage);
        } else if (smb == SMB_ROUNDOFF) {
            if (alpha) launch_stage_lo<SMB_ROUNDOFF, true>(P, U, Uo, out, d0, stage);
            else launch_stage_lo<SMB_ROUNDOFF, false>(P, U, Uo, out, d0, stage);
        } else {
            if (alpha) launch_stage_lo<SMB_2, true>(P, U, Uo, out, d0, stage);
            else launch_stage_lo<SMB_2, false>(P, U, Uo, out, d0, stage);
        }
        timed_end(P, KID_STAGE_LO);
    } else if (md.viscosity_type == SWE_FCT) {
        timed_begin(P, KID_FCT_LO);
        if (alpha) launch_fct_lo<true, true>(P, U, d0);
        else launch_fct_lo<false, true>(P, U, d0);
        timed_end(P, KID_FCT_LO);
        {
            const LaunchCfg& L = P.tune.k[KID_RMINUS];
            timed_begin(P, KID_RMINUS);
            SWE_BS_SWITCH(L.block, hipLaunchKernelGGL((k_fct_rminus<BS>), dim3(blocks_of((int64_t)M.nact * 64, BS)), dim3(BS), 0,
                                                      P.stream, M, T, U, (const double*)P.ulo_h, (const double2*)P.dmu,
                                                      (const double*)P.res, P.Rm, L.remap))
            timed_end(P, KID_RMINUS);
        }
        {
            const LaunchCfg& L = P.tune.k[KID_APPLY];
            timed_begin(P, KID_APPLY);
            SWE_BS_SWITCH(L.block, hipLaunchKernelGGL((k_fct_apply<BS>), dim3(blocks_of((int64_t)M.nact * 64, BS)), dim3(BS), 0,
                                                      P.stream, M, T, U, Uo, (const double*)P.ulo_h, (const double2*)P.ulo_q, out,
                                                      (const double2*)P.dmu, (const double*)P.res,
                                                      (const double*)P.Rm, stage, P.flags, L.remap))
            timed_end(P, KID_APPLY);
        }
    } else {  // entropy_visc
        if (alpha) launch_fct_lo<true, false>(P, U, d0);
        else launch_fct_lo<false, false>(P, U, d0);
        const LaunchCfg& L = P.tune.k[KID_EV];
        timed_begin(P, KID_EV);
        constexpr int BS = 256;
        const dim3 g(blocks_of((int64_t)M.nact * 64, BS));
        if (P.any_src)
            hipLaunchKernelGGL((k_ev_update<BS, true>), g, dim3(BS), 0, P.stream, M, T, U, Uo, out,
                               (const double2*)P.dmu, (const double*)P.res, md.roundoff, stage,
                               P.flags, L.remap);
        else
            hipLaunchKernelGGL((k_ev_update<BS, false>), g, dim3(BS), 0, P.stream, M, T, U, Uo, out,
                               (const double2*)P.dmu, (const double*)P.res, md.roundoff, stage,
                               P.flags, L.remap);
        timed_end(P, KID_EV);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SWE_EHIP, "kernel launch failed: %s", hipGetErrorString(e));
    return SWE_OK;
}

void launch_bdy(Pool& P, Node* u, int stage) {
    const int64_t n = P.nbc();
    if (n == 0) return;
    hipLaunchKernelGGL(k_bdy, dim3(blocks_of(n, 256)), dim3(256), 0, P.stream, u, (const BcEntry*)P.bc, (int)n,
                       (const double*)P.bcv, P.bcv_stride, (const InstDev*)P.tab, stage);
}

bool persistent_applies(const Pool& P) {
    return P.mode_set && P.mode.viscosity_type == SWE_FCT && !P.mode.sgn && P.coop_ok;
}

int launch_step_persistent(Pool& P, int nsteps) {
    const Mesh M = mesh_view(P);
    if (M.nact == 0 || nsteps <= 0) return SWE_OK;
    if (!persistent_applies(P)) return fail(SWE_EBADOPT, "the persistent step covers the fct scheme only");
    // SWE_BARRIER_SPINS (environment) lowers the barrier's wait bound, to exercise the path where a
    // barrier gives up (every workgroup leaves, the stepping instances fail)
    const char* sp = std::getenv("SWE_BARRIER_SPINS");
    const long spins = sp ? std::max(0L, std::atol(sp)) : kBarrierSpins;
    StepArgs A{M,       P.tab,    P.flags,  P.S[0],  P.S[1],        P.S[2],    P.d0,
               P.alpha, P.ulo_h,  P.ulo_q,  P.dmu,   P.res,         P.Rm,      P.vmax,
               P.bc,    P.bcv,    P.bcv_stride,      (int)P.nbc(),  P.ninst(), P.mode.roundoff, P.bar, spins};
    const size_t lds = (size_t)(2 + M.wmax) * 64 * sizeof(double) * 4;
    const bool alpha = P.mode.alpha != 0;
    const void* fn = alpha ? (P.any_src ? (const void*)k_step_fct<true, true> : (const void*)k_step_fct<true, false>)
                           : (P.any_src ? (const void*)k_step_fct<false, true> : (const void*)k_step_fct<false, false>);
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, lds);
    if (e != hipSuccess || per_cu < 1)
        return fail(SWE_EHIP, "persistent step: no workgroup fits a CU (%s)", hipGetErrorString(e));
    // every workgroup resident at once; no more than the slices need
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * P.num_cu, ((int64_t)M.nact + 3) / 4));
    // the arrival count stays below 2^31 within a launch: at most 17 barriers per step
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(256, ((int64_t)1 << 30) / (17 * (int64_t)blocks)));
    for (int done = 0; done < nsteps; done += chunk) {
        int n = std::min(chunk, nsteps - done);
        hipLaunchKernelGGL(k_poke, dim3(1), dim3(kPokeWords), 0, P.stream, (uint32_t*)P.bar, PokeWords{}, 2);
        void* args[] = {(void*)&A, (void*)&n};
        e = hipLaunchCooperativeKernel(fn, dim3(blocks), dim3(256), args, (unsigned)lds, P.stream);
        if (e != hipSuccess) return fail(SWE_EHIP, "persistent step launch: %s", hipGetErrorString(e));
        hipLaunchKernelGGL(k_step_check, dim3(blocks_of(P.ninst(), 64)), dim3(64), 0, P.stream, (const unsigned*)P.bar,
                           (const InstDev*)P.tab, P.flags, P.ninst());
    }
    e = hipGetLastError();
    if (e != hipSuccess) return fail(SWE_EHIP, "persistent step: %s", hipGetErrorString(e));
    return SWE_OK;
}

void launch_dij_full(Pool& P, const Node* U, double* dij) {
    const Mesh M = mesh_view(P);
    if (M.nact == 0) return;
    hipLaunchKernelGGL((k_dij_full<256>), dim3(blocks_of((int64_t)M.nact * 64, 256)), dim3(256), 0, P.stream, M,
                       (const InstDev*)P.tab, U, dij, P.flags);
}

void launch_pack5(Pool& P, const InstHost& I, const double* un5, Node* dst, NodeX* xdst) {
    const int64_t r0 = I.slice0 * 64;
    hipLaunchKernelGGL(k_pack5, dim3(blocks_of(I.np, 256)), dim3(256), 0, P.stream, un5, (const double*)(P.z + r0),
                       (const int*)(P.perm + r0), dst + r0, xdst + r0, I.np);
}

void launch_unpack5(Pool& P, const InstHost& I, const Node* src, const NodeX* xsrc, double* un5, hipStream_t stream) {
    const int64_t r0 = I.slice0 * 64;
    hipLaunchKernelGGL(k_unpack5, dim3(blocks_of(I.np, 256)), dim3(256), 0, stream ? stream : P.stream, src + r0,
                       xsrc + r0, (const int*)(P.perm + r0), un5, I.np);
}

void launch_pack(Pool& P, const InstHost& I, const double* un3, Node* dst) {
    const int64_t r0 = I.slice0 * 64;
    hipLaunchKernelGGL(k_pack, dim3(blocks_of(I.np, 256)), dim3(256), 0, P.stream, un3, (const double*)(P.z + r0),
                       (const int*)(P.perm + r0), dst + r0, I.np);
}

void launch_unpack(Pool& P, const InstHost& I, const Node* src, double* un3, hipStream_t stream) {
    const int64_t r0 = I.slice0 * 64;
    hipLaunchKernelGGL(k_unpack, dim3(blocks_of(I.np, 256)), dim3(256), 0, stream ? stream : P.stream, src + r0,
                       (const int*)(P.perm + r0), un3, I.np);
}

void launch_perm_nodes(Pool& P, const InstHost& I, const double* src, double* dst) {
    const int64_t r0 = I.slice0 * 64;
    hipLaunchKernelGGL(k_perm_nodes, dim3(blocks_of(I.np, 256)), dim3(256), 0, P.stream, src,
                       (const int*)(P.perm + r0), dst + r0, I.np);
}

void launch_csr_to_sell(Pool& P, const double* src, const int64_t* map, double* dst, int64_t nnz) {
    hipLaunchKernelGGL(k_csr_to_sell, dim3(blocks_of(nnz, 256)), dim3(256), 0, P.stream, src, map, dst, nnz);
}

void launch_sell_to_csr(Pool& P, const double* src, const int64_t* map, double* dst, int64_t nnz) {
    hipLaunchKernelGGL(k_sell_to_csr, dim3(blocks_of(nnz, 256)), dim3(256), 0, P.stream, src, map, dst, nnz);
}

void launch_fctpos(Pool& P, const double* f, const double* us, double htiny, double* lij) {
    const Mesh M = mesh_view(P);
    if (M.nact == 0) return;
    hipLaunchKernelGGL(k_fctpos_r, dim3(blocks_of((int64_t)M.nact * 64, 256)), dim3(256), 0, P.stream, M, f, us, htiny, P.Rm);
    hipLaunchKernelGGL(k_fctpos_l, dim3(blocks_of((int64_t)M.nact * 64, 256)), dim3(256), 0, P.stream, M, f, (const double*)P.Rm,
                       lij);
}

void launch_fill_slices(Pool& P, int64_t s0, int64_t n, int inst) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_slices, dim3(blocks_of(n, 256)), dim3(256), 0, P.stream, P.sl, s0, n, inst);
}

int launch_poke(Pool& P, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SWE_OK;
    if (bytes % 4 || bytes > sizeof(PokeWords) || ((uintptr_t)dst & 3))
        return fail(SWE_EINVAL, "table edit of %zu bytes is not a whole number of aligned words", bytes);
    PokeWords w;
    std::memcpy(w.w, src, bytes);
    hipLaunchKernelGGL(k_poke, dim3(1), dim3(kPokeWords), 0, P.stream, (uint32_t*)dst, w, (int)(bytes / 4));
    SWE_HIPCHK(hipGetLastError());
    return SWE_OK;
}

void launch_fill_bc(Pool& P, int64_t b0, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_bc, dim3(blocks_of(n, 256)), dim3(256), 0, P.stream, P.bc, b0, n);
}

int launch_fctpos_csr(int np, int nnz, const int* ia, const int* ja, const double* f, const double* us,
                      double htiny, double* lij) {
    double* dR = nullptr;
    hipError_t e = hipMalloc((void**)&dR, sizeof(double) * (size_t)np);
    if (e != hipSuccess) return fail(SWE_ENOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
    const unsigned g = blocks_of(np, 256);
    hipLaunchKernelGGL(k_fctpos_csr_r, dim3(g), dim3(256), 0, 0, np, nnz, ia, f, us, dR, htiny);
    hipLaunchKernelGGL(k_fctpos_csr_l, dim3(g), dim3(256), 0, 0, np, nnz, ia, ja, f, (const double*)dR, lij);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(dR);
    if (e != hipSuccess) return fail(SWE_EHIP, "FCT_positivity kernels: %s", hipGetErrorString(e));
    return SWE_OK;
}

int launch_fctgen_csr(int np, const int* ia, const int* ja, const double* f, const double* maxn, const double* minn,
                      const double* dg, double* u) {
    double2* dR = nullptr;
    hipError_t e = hipMalloc((void**)&dR, sizeof(double2) * (size_t)np);
    if (e != hipSuccess) return fail(SWE_ENOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
    const unsigned g = blocks_of(np, 256);
    // every R+- is complete before any row applies its fluxes: two launches on one stream
    hipLaunchKernelGGL(k_fctgen_csr_r, dim3(g), dim3(256), 0, 0, np, ia, f, (const double*)u, maxn, minn, dg, dR);
    hipLaunchKernelGGL(k_fctgen_csr_apply, dim3(g), dim3(256), 0, 0, np, ia, ja, f, (const double2*)dR, dg, u);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(dR);
    if (e != hipSuccess) return fail(SWE_EHIP, "FCT_generic kernels: %s", hipGetErrorString(e));
    return SWE_OK;
}

// ---- HBM streaming probes -------------------------------------------------------------------------
// The achievable bandwidth of the device, the second ceiling the step's kernels are quoted against:
// a streaming read (16-byte non-temporal loads, grid-stride, a per-block sum written out so the loads
// stay live) and a streaming copy (16-byte non-temporal loads and stores).
namespace {
__global__ void __launch_bounds__(256) k_stream_read(const dvec2* __restrict__ a, size_t n2, double* __restrict__ out) {
    double s = 0.0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += stride) {
        const dvec2 v = __builtin_nontemporal_load(a + q);
        s += v.x + v.y;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_stream_copy(const dvec2* __restrict__ a, dvec2* __restrict__ b, size_t n2) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n2; q += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(a + q), b + q);
}
}  // namespace

int stream_bandwidth(size_t bytes, int reps, double* read_gbs, double* copy_gbs) {
    bytes = bytes / 16 * 16;
    if (bytes == 0 || reps < 1) return fail(SWE_EINVAL, "stream probe: empty size or no repetitions");
    const size_t n2 = bytes / 16;
    const int blocks = 8192;
    dvec2 *a = nullptr, *b = nullptr;
    double* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc((void**)&a, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&b, bytes);
    if (e == hipSuccess) e = hipMalloc((void**)&out, sizeof(double) * blocks);
    if (e == hipSuccess) e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best_r = 1e30f, best_c = 1e30f;
    for (int r = 0; r <= reps && e == hipSuccess; ++r) {     // the first pass warms up
        float ms = 0.0f;
        hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(256), 0, 0, (const dvec2*)a, n2, out);
        e = hipEventRecord(e0, 0);
        hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(256), 0, 0, (const dvec2*)a, n2, out);
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) best_r = std::min(best_r, ms);
        if (e == hipSuccess) e = hipEventRecord(e0, 0);
        hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, 0, (const dvec2*)a, b, n2);
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (r > 0) best_c = std::min(best_c, ms);
    }
    if (e == hipSuccess) e = hipGetLastError();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return fail(SWE_EHIP, "stream probe: %s", hipGetErrorString(e));
    *read_gbs = (double)bytes / (best_r * 1e-3) / 1e9;
    *copy_gbs = 2.0 * (double)bytes / (best_c * 1e-3) / 1e9;
    return SWE_OK;
}

// ---- FCT_positivity without allocations -----------------------------------------------------
// R- of every row lives in `scratch`, a fixed array in the code object: each GPU the library runs
// on holds its own copy (kNoAllocRows doubles, 32 MiB) until the process ends, and every call on
// that GPU uses it, so calls that overlap on one GPU (two streams, two host threads) race.  Each
// call writes R- of all its rows before any row reads it, so nothing carries over between calls.
// In host code `scratch` names a host-side stand-in, not device memory: kernels get the device
// address that hipGetSymbolAddress returns, looked up once per device on the first call there.
static __device__ __align__(256) double scratch[kNoAllocRows];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_of_current_device(double** out) {
    int dev = -1;
    SWE_HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(SWE_EINVAL, "device %d is beyond the %d this library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(SWE_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return SWE_OK;
}
}  // namespace

int launch_fctpos_csr_noalloc(int np, int nnz, const int* ia, const int* ja, const double* f, const double* us,
                              double htiny, double* lij, hipStream_t stream) {
    double* R = nullptr;
    int rc = scratch_of_current_device(&R);
    if (rc) return rc;
    const unsigned g = blocks_of(np, 256);
    // two launches on one stream: every R- is written before any row forms its l_ij
    hipLaunchKernelGGL(k_fctpos_csr_r, dim3(g), dim3(256), 0, stream, np, nnz, ia, f, us, R, htiny);
    hipLaunchKernelGGL(k_fctpos_csr_l, dim3(g), dim3(256), 0, stream, np, nnz, ia, ja, f, (const double*)R, lij);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SWE_EHIP, "FCT_positivity kernels: %s", hipGetErrorString(e));
    return SWE_OK;
}

}  // namespace swe
