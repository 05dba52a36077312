// swe_capi.cpp -- the C-ABI of include/swe_amd.h over the pool engine.
//
// A swe_handle is a pool sized for exactly one mesh; its entry points follow the reference's
// routines (update.f90, fct.f90, main.f90) on host arrays or on the device-resident state.  A
// swe_batch is a pool sized by the caller that admits and releases meshes while it steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstddef>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "swe_amd.h"
#include "swe_engine.h"

using namespace swe;

struct swe_handle {
    Pool* P = nullptr;
    int np = 0, nnz = 0;
    int64_t* csr2slot = nullptr;    // caller's CSR entry -> slot
    double* dt_log = nullptr;       // ring of dt per step
    int dt_cap = 0;
    long long queued = 0;           // steps queued so far (the instance's step counter)
    bool state_set = false;
    bool step_open = false;         // swe_step_begin ran; swe_step_finish completes the step
    int bc_kind = SWE_BC_NONE;
};

struct swe_batch {
    Pool* P = nullptr;
};

struct swe_errors {
    swe_handle* h = nullptr;     // a solver's evaluator, or
    swe::Pool* P = nullptr;      // a batch instance's: the pool, the id and its admission's ticket
    int k = 0;
    long long since = 0;
    swe::ErrEval* E = nullptr;
};

namespace {

constexpr int kId = 0;   // the handle's instance

int check_flags(Pool* P, int k) {
    int f = 0;
    SWE_HIPCHK(hipMemcpyAsync(&f, P->flags + k, sizeof(int), hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    if (f) {
        // report once; the instance may step again afterwards, as the reference run would not
        const int z = 0, st = INST_RUNNING;
        int rc = launch_poke(*P, P->flags + k, &z, sizeof z);
        if (!rc) rc = launch_poke(*P, &P->tab[k].state, &st, sizeof st);
        if (rc) return rc;
        SWE_HIPCHK(hipStreamSynchronize(P->stream));
        if (f & FLAG_STEP) return fail(SWE_EHIP, "persistent step: a grid barrier gave up waiting");
        if (f & FLAG_LAMBDA) return fail(SWE_ELAMBDA, "BUG in compute_lambda_vacc (Delta < 0)");
        if (f & FLAG_NEG_H) return fail(SWE_ENEG_H, "Min h<0, STOP");
    }
    return SWE_OK;
}

int read_field(Pool* P, int k, size_t off, void* dst, size_t bytes) {
    SWE_HIPCHK(hipMemcpyAsync(dst, (const char*)(P->tab + k) + off, bytes, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    return SWE_OK;
}

void free_handle(swe_handle* h) {
    if (!h) return;
    if (h->csr2slot) (void)hipFree(h->csr2slot);
    if (h->dt_log) (void)hipFree(h->dt_log);
    pool_destroy(h->P);
    delete h;
}

int ensure_dt_log(swe_handle* h, int n) {
    if (n <= h->dt_cap) return SWE_OK;
    SWE_HIPCHK(hipStreamSynchronize(h->P->stream));
    if (h->dt_log) (void)hipFree(h->dt_log);
    h->dt_log = nullptr;
    h->dt_cap = 0;
    const int cap = std::max(n, 4096);
    const hipError_t e = hipMalloc((void**)&h->dt_log, sizeof(double) * (size_t)cap);
    if (e != hipSuccess) return fail(SWE_ENOMEM, "hipMalloc(dt log): %s", hipGetErrorString(e));
    h->dt_cap = cap;
    int rc = pool_set_field(h->P, kId, offsetof(InstDev, dt_log), &h->dt_log, sizeof(double*));
    if (rc) return rc;
    return pool_set_field(h->P, kId, offsetof(InstDev, dt_cap), &h->dt_cap, sizeof(int));
}

int timing_enable(Pool* P, const char* kernel) {
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    P->timing = (kernel != nullptr && kernel[0] != 0);
    P->timed_name = P->timing ? kernel : "";
    P->tk.used = 0;
    return SWE_OK;
}

int timing_read(Pool* P, const char* kernel, int* launches, double* ms) {
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    *launches = 0;
    *ms = 0.0;
    if (P->timed_name != kernel) return SWE_OK;
    double sum = 0.0;
    for (int k = 0; k < P->tk.used; ++k) {
        float t = 0.0f;
        SWE_HIPCHK(hipEventElapsedTime(&t, P->tk.ev[(size_t)2 * k], P->tk.ev[(size_t)2 * k + 1]));
        sum += t;
    }
    *launches = P->tk.used;
    *ms = sum;
    return SWE_OK;
}

int set_bc_values(Pool* P, int k, const double* hv, const double* uxv, const double* uyv) {
    const InstHost& I = P->inst[(size_t)k];
    if ((I.nh && !hv) || (I.nux && !uxv) || (I.nuy && !uyv)) return fail(SWE_EINVAL, "missing bc values");
    // values follow the entry order of admission: h list, ux list, uy list
    if (I.nh) SWE_HIPCHK(hipMemcpyAsync(P->bcv + I.bc0, hv, sizeof(double) * I.nh, hipMemcpyHostToDevice, P->stream));
    if (I.nux)
        SWE_HIPCHK(hipMemcpyAsync(P->bcv + I.bc0 + I.nh, uxv, sizeof(double) * I.nux, hipMemcpyHostToDevice, P->stream));
    if (I.nuy)
        SWE_HIPCHK(hipMemcpyAsync(P->bcv + I.bc0 + I.nh + I.nux, uyv, sizeof(double) * I.nuy, hipMemcpyHostToDevice,
                                  P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    return SWE_OK;
}

int get_schedule(Pool* P, int* blocks, int* remap) {
    for (int k = 0; k < KID_COUNT; ++k) {
        if (blocks) blocks[k] = P->tune.k[k].block;
        if (remap) remap[k] = P->tune.k[k].remap;
    }
    return SWE_OK;
}

int get_schedule_kinds(Pool* P, int* kinds) {
    for (int k = 0; k < KID_COUNT; ++k) kinds[k] = P->tune.k[k].kind;
    return SWE_OK;
}

int set_schedule_kinds(Pool* P, const int* kinds) {
    Tuning t = P->tune;
    for (int k = 0; k < KID_COUNT; ++k) {
        if (kinds[k] < 0 || kinds[k] >= kinds_of(k))
            return fail(SWE_EINVAL, "%s has no schedule kind %d", kKernelNames[k], kinds[k]);
        t.k[k].kind = kinds[k];
    }
    P->tune = t;
    return SWE_OK;
}

int set_schedule(Pool* P, const int* blocks, const int* remap) {
    Tuning t = P->tune;
    for (int k = 0; k < KID_COUNT; ++k) {
        if (blocks) {
            if (!block_supported(k, blocks[k])) return fail(SWE_EINVAL, "unsupported workgroup size %d", blocks[k]);
            t.k[k].block = blocks[k];
        }
        if (remap) t.k[k].remap = remap[k] ? 1 : 0;
    }
    P->tune = t;
    return SWE_OK;
}

bool live(const Pool* P, int id) { return id >= 0 && id < P->inst_cap && P->inst[(size_t)id].used; }

}  // namespace

// =============================================================================================
extern "C" {

const char* swe_last_error(void) { return last_error(); }

int swe_create(const swe_mesh_desc* d, const swe_params* prm, swe_handle** out) {
    clear_error();
    if (!d || !prm || !out) return fail(SWE_EINVAL, "null argument");
    *out = nullptr;
    int rc;
    HostLayout L;
    Mode md;
    if ((rc = mode_of(*prm, md))) return rc;
    if ((rc = check_sgn_lists(md, d->n_ux_dir, d->n_uy_dir))) return rc;
    if ((rc = build_layout(*d, sort_window_default(), L))) return rc;
    PoolCaps caps;
    caps.rows = L.np_pad;
    caps.slots = L.nslot;
    caps.elist = (int64_t)L.elist.size();
    caps.bc = (int64_t)(L.lists[0].size() + L.lists[1].size() + L.lists[2].size());
    caps.instances = 1;
    swe_handle* h = new swe_handle();
    h->np = d->np;
    h->nnz = d->nnz;
    if ((rc = pool_create(caps, &h->P))) {
        delete h;
        return rc;
    }
    h->P->use_graph = true;
    int id = -1;
    if ((rc = pool_add(h->P, *d, *prm, nullptr, 0.0, std::numeric_limits<double>::infinity(), LLONG_MAX, &id, &L))) {
        free_handle(h);
        return rc;
    }
    hipError_t e = hipMalloc((void**)&h->csr2slot, sizeof(int64_t) * (size_t)d->nnz);
    if (e == hipSuccess)
        e = hipMemcpy(h->csr2slot, L.csr2slot.data(), sizeof(int64_t) * (size_t)d->nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamSynchronize(h->P->stream);
    if (e != hipSuccess) {
        free_handle(h);
        return fail(SWE_EHIP, "swe_create: %s", hipGetErrorString(e));
    }
    *out = h;
    return SWE_OK;
}

int swe_create_from_assembly(const swe_assembly* a, const double* bath, int n_h_dir, const int* h_dir,
                             int n_ux_dir, const int* ux_dir, int n_uy_dir, const int* uy_dir, const swe_params* prm,
                             swe_handle** out) {
    clear_error();
    if (!a || !bath || !prm || !out) return fail(SWE_EINVAL, "null argument");
    *out = nullptr;
    int rc;
    Mode md;
    if ((rc = mode_of(*prm, md))) return rc;
    if ((rc = check_sgn_lists(md, n_ux_dir, n_uy_dir))) return rc;
    DeviceLayout DL;
    if ((rc = build_layout_device(a, DL))) return rc;
    if (DL.nnz > INT_MAX) {
        free_device_layout(DL);
        return fail(SWE_EINVAL, "%lld entries do not fit the solver's 32-bit CSR", (long long)DL.nnz);
    }
    const int* lists[3] = {h_dir, ux_dir, uy_dir};
    const int nlist[3] = {n_h_dir, n_ux_dir, n_uy_dir};
    PoolCaps caps;
    caps.rows = (int64_t)DL.nsl * 64;
    caps.slots = DL.nslot;
    caps.elist = DL.ne;
    caps.bc = (int64_t)std::max(0, n_h_dir) + std::max(0, n_ux_dir) + std::max(0, n_uy_dir);
    caps.instances = 1;
    swe_handle* h = new swe_handle();
    h->np = DL.np;
    h->nnz = (int)DL.nnz;
    if ((rc = pool_create(caps, &h->P))) {
        delete h;
        free_device_layout(DL);
        return rc;
    }
    h->P->use_graph = true;
    hipError_t e = hipMalloc((void**)&h->csr2slot, sizeof(int64_t) * (size_t)std::max<int64_t>(1, DL.nnz));
    if (e != hipSuccess) {
        free_handle(h);
        free_device_layout(DL);
        return fail(SWE_EHIP, "swe_create_from_assembly: %s", hipGetErrorString(e));
    }
    int id = -1;
    rc = pool_add_assembled(h->P, a, DL, bath, lists, nlist, *prm, nullptr, 0.0,
                            std::numeric_limits<double>::infinity(), LLONG_MAX, &id, h->csr2slot);
    free_device_layout(DL);
    if (rc) {
        free_handle(h);
        return rc;
    }
    *out = h;
    return SWE_OK;
}

int swe_errors_create(swe_handle* h, int me, const int* jj, const double* rr, swe_errors** out) {
    clear_error();
    if (!h || !jj || !rr || !out) return fail(SWE_EINVAL, "null argument");
    *out = nullptr;
    swe::ErrEval* E = nullptr;
    const int rc = errors_create(h->P, kId, me, jj, rr, &E);
    if (rc) return rc;
    *out = new swe_errors{h, h->P, kId, h->P->inst[(size_t)kId].since, E};
    return SWE_OK;
}

int swe_compute_errors(swe_errors* e, double time, double* out) {
    if (!e || !out) return fail(SWE_EINVAL, "null argument");
    if (swe_handle* h = e->h) {
        if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
        if (h->step_open) return fail(SWE_EINVAL, "a step is open (swe_step_begin without swe_step_finish)");
    }
    Pool* P = e->P;
    const InstHost& I = P->inst[(size_t)e->k];
    if (!I.used || I.since != e->since) return fail(SWE_EINVAL, "instance %d is no longer the evaluator's mesh", e->k);
    return errors_eval(P, e->E, P->S[0], I.prm.type_test, time, I.prm.gravity, I.prm.htiny, out);
}

int swe_ns_l1_ratio(swe_errors* e, double* out) {
    if (!e || !out) return fail(SWE_EINVAL, "null argument");
    if (swe_handle* h = e->h) {
        if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
        if (h->step_open) return fail(SWE_EINVAL, "a step is open (swe_step_begin without swe_step_finish)");
    }
    Pool* P = e->P;
    const InstHost& I = P->inst[(size_t)e->k];
    if (!I.used || I.since != e->since) return fail(SWE_EINVAL, "instance %d is no longer the evaluator's mesh", e->k);
    return errors_flow(P, e->E, P->S[0], I.prm.gravity, out);
}

int swe_errors_destroy(swe_errors* e) {
    if (!e) return SWE_OK;
    errors_destroy(e->E);
    delete e;
    return SWE_OK;
}

int swe_destroy(swe_handle* h) {
    if (!h) return SWE_OK;
    free_handle(h);
    return SWE_OK;
}

int swe_set_params(swe_handle* h, const swe_params* prm) {
    if (!h || !prm) return fail(SWE_EINVAL, "null argument");
    return pool_set_params(h->P, kId, *prm);
}

int swe_set_bc(swe_handle* h, int kind, double amp, double period, double level) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    const int rc = pool_set_bc(h->P, kId, kind, amp, period, level);
    if (!rc) h->bc_kind = kind;
    return rc;
}

int swe_set_bc_values(swe_handle* h, const double* hv, const double* uxv, const double* uyv) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    return set_bc_values(h->P, kId, hv, uxv, uyv);
}

int swe_set_state(swe_handle* h, const double* un) {
    if (!h || !un) return fail(SWE_EINVAL, "null argument");
    int rc = pool_upload_state(h->P, kId, un, h->P->S[0]);
    if (rc) return rc;
    SWE_HIPCHK(hipStreamSynchronize(h->P->stream));
    h->state_set = true;
    return SWE_OK;
}

int swe_get_state(swe_handle* h, double* un) {
    if (!h || !un) return fail(SWE_EINVAL, "null argument");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    return pool_download_state(h->P, kId, h->P->S[0], un);
}

int swe_compute_dij(swe_handle* h, const double* un, double* dij) {
    if (!h || !un || !dij) return fail(SWE_EINVAL, "null argument");
    Pool* P = h->P;
    int rc = pool_upload_state(P, kId, un, P->scratch);
    if (rc) return rc;
    launch_dij_full(*P, P->scratch, P->dL());
    launch_sell_to_csr(*P, P->dL(), h->csr2slot, P->ework, h->nnz);
    SWE_HIPCHK(hipMemcpyAsync(dij, P->ework, sizeof(double) * (size_t)h->nnz, hipMemcpyDeviceToHost, P->stream));
    return check_flags(P, kId);
}

int swe_compute_dt(swe_handle* h, const double* un, double* dt) {
    if (!h || !dt) return fail(SWE_EINVAL, "null argument");
    Pool* P = h->P;
    const Node* U = P->S[0];
    if (un) {
        int rc = pool_upload_state(P, kId, un, P->scratch);
        if (rc) return rc;
        U = P->scratch;
    } else if (!h->state_set) {
        return fail(SWE_EINVAL, "no state has been set");
    }
    launch_compute_dt(*P, U, nullptr, 1);
    launch_finalize_dt(*P, 0);
    int rc = check_flags(P, kId);
    const int rc2 = read_field(P, kId, offsetof(InstDev, dt), dt, sizeof(double));
    return rc ? rc : rc2;
}

int swe_euler(swe_handle* h, const double* un, double* unext, double dt) {
    if (!h || !un || !unext) return fail(SWE_EINVAL, "null argument");
    Pool* P = h->P;
    int rc = pool_upload_state(P, kId, un, P->scratch);
    if (rc) return rc;
    const int go = 1;
    if ((rc = pool_set_field(P, kId, offsetof(InstDev, dt), &dt, sizeof dt))) return rc;
    if ((rc = pool_set_field(P, kId, offsetof(InstDev, go), &go, sizeof go))) return rc;
    if ((rc = launch_euler_stage(*P, P->scratch, P->scratch, P->S[2], nullptr, 0))) return rc;
    if ((rc = pool_download_state(P, kId, P->S[2], unext))) return rc;
    return check_flags(P, kId);
}

int swe_fct_positivity(swe_handle* h, const double* us, const double* f, double* lij) {
    if (!h || !us || !f || !lij) return fail(SWE_EINVAL, "null argument");
    Pool* P = h->P;
    const InstHost& I = P->inst[kId];
    double htiny = 0.0;
    int rc = read_field(P, kId, offsetof(InstDev, K) + offsetof(Consts, htiny), &htiny, sizeof htiny);
    if (rc) return rc;
    SWE_HIPCHK(hipMemcpyAsync(P->stage3, us, sizeof(double) * (size_t)h->np, hipMemcpyHostToDevice, P->stream));
    SWE_HIPCHK(hipMemcpyAsync(P->ework, f, sizeof(double) * (size_t)h->nnz, hipMemcpyHostToDevice, P->stream));
    launch_perm_nodes(*P, I, P->stage3, P->nwork);
    launch_csr_to_sell(*P, P->ework, h->csr2slot, P->d0, h->nnz);
    launch_fctpos(*P, P->d0, P->nwork, htiny, P->dL());
    launch_sell_to_csr(*P, P->dL(), h->csr2slot, P->ework, h->nnz);
    SWE_HIPCHK(hipMemcpyAsync(lij, P->ework, sizeof(double) * (size_t)h->nnz, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
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
    double *dus = nullptr, *df = nullptr, *dl = nullptr;
    hipError_t e = hipMalloc((void**)&dia, sizeof(int) * ((size_t)np + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&dja, sizeof(int) * (size_t)nnz);
    if (e == hipSuccess) e = hipMalloc((void**)&dus, sizeof(double) * (size_t)np);
    if (e == hipSuccess) e = hipMalloc((void**)&df, sizeof(double) * (size_t)nnz);
    if (e == hipSuccess) e = hipMalloc((void**)&dl, sizeof(double) * (size_t)nnz);
    int rc = SWE_OK;
    if (e != hipSuccess) rc = fail(SWE_ENOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
    if (!rc) {
        e = hipMemcpy(dia, ia, sizeof(int) * ((size_t)np + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dja, ja, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dus, us, sizeof(double) * (size_t)np, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(df, f, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(SWE_EHIP, "swe_fct_positivity_csr: %s", hipGetErrorString(e));
    }
    if (!rc) rc = launch_fctpos_csr(np, nnz, dia, dja, df, dus, htiny, dl);
    if (!rc) {
        e = hipMemcpy(lij, dl, sizeof(double) * (size_t)nnz, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(SWE_EHIP, "swe_fct_positivity_csr: %s", hipGetErrorString(e));
    }
    void* bufs[] = {dia, dja, dus, df, dl};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    return rc;
}

int swe_fct_positivity_csr_noalloc(int np, int nnz, const int* ia, const int* ja, const double* us, const double* f,
                                   double htiny, double* lij, void* stream) {
    clear_error();
    if (np <= 0 || nnz <= 0 || !ia || !ja || !us || !f || !lij) return fail(SWE_EINVAL, "null or empty argument");
    if (np > kNoAllocRows)
        return fail(SWE_EINVAL, "swe_fct_positivity_csr_noalloc: %d rows exceed its fixed scratch of %d rows", np,
                    kNoAllocRows);
    return launch_fctpos_csr_noalloc(np, nnz, ia, ja, f, us, htiny, lij, (hipStream_t)stream);
}

int swe_stream_bandwidth(unsigned long long bytes, int reps, double* read_gbs, double* copy_gbs) {
    clear_error();
    if (!read_gbs || !copy_gbs) return fail(SWE_EINVAL, "null argument");
    return stream_bandwidth((size_t)bytes, reps, read_gbs, copy_gbs);
}

int swe_fct_generic_csr(int np, int nnz, const int* ia, const int* ja, double* unext, const double* maxn,
                        const double* minn, const double* f, const double* dg) {
    if (np <= 0 || nnz <= 0 || !ia || !ja || !unext || !maxn || !minn || !f || !dg)
        return fail(SWE_EINVAL, "null or empty argument");
    if (ia[0] != 1 || ia[np] - 1 != nnz) return fail(SWE_EINVAL, "ia must start at 1 and end at nnz+1");
    for (int i = 0; i < np; ++i)
        if (ia[i + 1] < ia[i]) return fail(SWE_EINVAL, "ia is not monotone at row %d", i + 1);
    for (int p = 0; p < nnz; ++p)
        if (ja[p] < 1 || ja[p] > np) return fail(SWE_EINVAL, "ja(%d)=%d out of range", p + 1, ja[p]);
    int *dia = nullptr, *dja = nullptr;
    double *du = nullptr, *dmax = nullptr, *dmin = nullptr, *df = nullptr, *ddg = nullptr;
    const size_t n = (size_t)np, z = (size_t)nnz;
    hipError_t e = hipMalloc((void**)&dia, sizeof(int) * (n + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&dja, sizeof(int) * z);
    if (e == hipSuccess) e = hipMalloc((void**)&du, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dmax, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&dmin, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&ddg, sizeof(double) * n);
    if (e == hipSuccess) e = hipMalloc((void**)&df, sizeof(double) * z);
    int rc = SWE_OK;
    if (e != hipSuccess) rc = fail(SWE_ENOMEM, "hipMalloc failed: %s", hipGetErrorString(e));
    if (!rc) {
        e = hipMemcpy(dia, ia, sizeof(int) * (n + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dja, ja, sizeof(int) * z, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(du, unext, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dmax, maxn, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dmin, minn, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(ddg, dg, sizeof(double) * n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(df, f, sizeof(double) * z, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = fail(SWE_EHIP, "swe_fct_generic_csr: %s", hipGetErrorString(e));
    }
    if (!rc) rc = launch_fctgen_csr(np, dia, dja, df, dmax, dmin, ddg, du);
    if (!rc) {
        e = hipMemcpy(unext, du, sizeof(double) * n, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(SWE_EHIP, "swe_fct_generic_csr: %s", hipGetErrorString(e));
    }
    void* bufs[] = {dia, dja, du, dmax, dmin, ddg, df};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    return rc;
}

int swe_set_time(swe_handle* h, double t) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    int rc = pool_set_field(h->P, kId, offsetof(InstDev, time), &t, sizeof t);
    if (rc) return rc;
    SWE_HIPCHK(hipStreamSynchronize(h->P->stream));
    return SWE_OK;
}

int swe_get_time(swe_handle* h, double* t) {
    if (!h || !t) return fail(SWE_EINVAL, "null argument");
    return read_field(h->P, kId, offsetof(InstDev, time), t, sizeof(double));
}

int swe_run(swe_handle* h, int nsteps, double* time, double* dts) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    int rc;
    if (time && (rc = swe_set_time(h, *time))) return rc;
    if (dts && (rc = ensure_dt_log(h, nsteps))) return rc;
    long long first = 0;
    if ((rc = read_field(h->P, kId, offsetof(InstDev, steps), &first, sizeof first))) return rc;
    if ((rc = pool_queue_steps(h->P, nsteps))) return rc;
    h->queued += nsteps;
    if (dts && nsteps > 0) {
        // ring positions first .. first+nsteps-1 (nsteps <= dt_cap)
        const long long a = first % h->dt_cap;
        const long long n1 = std::min<long long>(nsteps, h->dt_cap - a);
        SWE_HIPCHK(hipMemcpyAsync(dts, h->dt_log + a, sizeof(double) * (size_t)n1, hipMemcpyDeviceToHost, h->P->stream));
        if (n1 < nsteps)
            SWE_HIPCHK(hipMemcpyAsync(dts + n1, h->dt_log, sizeof(double) * (size_t)(nsteps - n1), hipMemcpyDeviceToHost,
                                      h->P->stream));
    }
    const int frc = check_flags(h->P, kId);
    if (time && (rc = swe_get_time(h, time))) return rc;
    return frc;
}

int swe_step_begin(swe_handle* h, double* dt) {
    if (!h || !dt) return fail(SWE_EINVAL, "null argument");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    if (h->step_open) return fail(SWE_EINVAL, "the step begun before has not been finished");
    Pool* P = h->P;
    int rc = record_dt(P);
    if (rc) return rc;
    SWE_HIPCHK(hipGetLastError());
    int go = 0;
    if ((rc = read_field(P, kId, offsetof(InstDev, go), &go, sizeof go))) return rc;
    if (!go) {
        // the instance takes no step: a check tripped earlier
        rc = check_flags(P, kId);
        return rc ? rc : fail(SWE_EINVAL, "the device-resident state takes no step");
    }
    if ((rc = read_field(P, kId, offsetof(InstDev, dt), dt, sizeof *dt))) return rc;
    h->step_open = true;
    h->queued += 1;
    return SWE_OK;
}

int swe_step_finish(swe_handle* h, const double* vals) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if (!h->step_open) return fail(SWE_EINVAL, "swe_step_begin has not been called");
    h->step_open = false;
    Pool* P = h->P;
    const InstHost& I = P->inst[(size_t)kId];
    if (h->bc_kind == SWE_BC_STAGES) {
        const int64_t n = (int64_t)I.nh + I.nux + I.nuy;
        if (n > 0 && !vals) return fail(SWE_EINVAL, "missing per-stage boundary values");
        for (int st = 0; st < 3 && n > 0; ++st)
            SWE_HIPCHK(hipMemcpyAsync(P->bcv + st * P->bcv_stride + I.bc0, vals + st * n, sizeof(double) * (size_t)n,
                                      hipMemcpyHostToDevice, P->stream));
    }
    int rc = record_stages(P);
    if (rc) return rc;
    SWE_HIPCHK(hipGetLastError());
    return check_flags(P, kId);
}

int swe_run_async(swe_handle* h, int nsteps) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    const int rc = pool_queue_steps(h->P, nsteps);
    if (!rc) h->queued += nsteps;
    return rc;
}

int swe_sync(swe_handle* h) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    return check_flags(h->P, kId);
}

void* swe_stream(swe_handle* h) { return h ? (void*)h->P->stream : nullptr; }

int swe_enable_kernel_timing(swe_handle* h, const char* kernel) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    return timing_enable(h->P, kernel);
}

int swe_kernel_timing(swe_handle* h, const char* kernel, int* launches, double* ms) {
    if (!h || !kernel || !launches || !ms) return fail(SWE_EINVAL, "null argument");
    return timing_read(h->P, kernel, launches, ms);
}

int swe_tune(swe_handle* h, int steps) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    if (!h->state_set) return fail(SWE_EINVAL, "no state has been set");
    return pool_tune(h->P, steps, 2);
}

int swe_get_schedule(swe_handle* h, int* blocks, int* remap) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    return get_schedule(h->P, blocks, remap);
}

int swe_set_schedule(swe_handle* h, const int* blocks, const int* remap) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    return set_schedule(h->P, blocks, remap);
}

int swe_get_schedule_kinds(swe_handle* h, int* kinds) {
    if (!h || !kinds) return fail(SWE_EINVAL, "null argument");
    return get_schedule_kinds(h->P, kinds);
}

int swe_set_schedule_kinds(swe_handle* h, const int* kinds) {
    if (!h || !kinds) return fail(SWE_EINVAL, "null argument");
    return set_schedule_kinds(h->P, kinds);
}

static int set_step_form(Pool* P, int form) {
    if (form != SWE_STEP_LAUNCHES && form != SWE_STEP_PERSISTENT) return fail(SWE_EINVAL, "no step form %d", form);
    if (form == SWE_STEP_PERSISTENT && !P->coop_ok)
        return fail(SWE_EINVAL, "the device does not take cooperative launches");
    P->step_form = form;
    return SWE_OK;
}

int swe_set_step_form(swe_handle* h, int form) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    return set_step_form(h->P, form);
}

int swe_get_step_form(swe_handle* h, int* form) {
    if (!h || !form) return fail(SWE_EINVAL, "null argument");
    *form = h->P->step_form;
    return SWE_OK;
}

int swe_batch_set_step_form(swe_batch* b, int form) {
    if (!b) return fail(SWE_EINVAL, "null batch");
    return set_step_form(b->P, form);
}

int swe_batch_get_step_form(swe_batch* b, int* form) {
    if (!b || !form) return fail(SWE_EINVAL, "null argument");
    *form = b->P->step_form;
    return SWE_OK;
}

int swe_use_graph(swe_handle* h, int on) {
    if (!h) return fail(SWE_EINVAL, "null handle");
    h->P->use_graph = on != 0;
    return SWE_OK;
}

// ---- batches ----------------------------------------------------------------------------------
static PoolCaps caps_of(const swe_batch_caps* c) {
    PoolCaps p;
    p.rows = c->max_nodes;
    p.slots = c->max_entries;
    // the compact off-diagonal lists take fewer elements than the slots
    p.elist = c->max_entries;
    p.bc = c->max_boundary;
    p.instances = c->max_instances;
    return p;
}

unsigned long long swe_batch_bytes(const swe_batch_caps* caps) {
    if (!caps) return 0ull;
    return (unsigned long long)pool_bytes(caps_of(caps));
}

int swe_batch_create(const swe_batch_caps* caps, swe_batch** out) {
    clear_error();
    if (!caps || !out) return fail(SWE_EINVAL, "null argument");
    *out = nullptr;
    swe_batch* b = new swe_batch();
    const int rc = pool_create(caps_of(caps), &b->P);
    if (rc) {
        delete b;
        return rc;
    }
    b->P->use_graph = true;
    *out = b;
    return SWE_OK;
}

int swe_batch_destroy(swe_batch* b) {
    if (!b) return SWE_OK;
    pool_destroy(b->P);
    delete b;
    return SWE_OK;
}

int swe_batch_add(swe_batch* b, const swe_mesh_desc* d, const swe_params* prm, const double* un, double t0,
                  double tfinal, long long max_steps, int* id) {
    if (!b || !d || !prm || !un || !id) return fail(SWE_EINVAL, "null argument");
    if (max_steps < 0) return fail(SWE_EINVAL, "max_steps < 0");
    return pool_add(b->P, *d, *prm, un, t0, tfinal, max_steps, id, nullptr);
}

int swe_batch_add_assembled(swe_batch* b, const swe_assembly* a, const double* bath, int n_h_dir, const int* h_dir,
                            int n_ux_dir, const int* ux_dir, int n_uy_dir, const int* uy_dir, const swe_params* prm,
                            const double* un, double t0, double tfinal, long long max_steps, int* id) {
    clear_error();
    if (!b || !a || !bath || !prm || !un || !id) return fail(SWE_EINVAL, "null argument");
    if (max_steps < 0) return fail(SWE_EINVAL, "max_steps < 0");
    DeviceLayout DL;
    int rc = build_layout_device(a, DL);
    if (rc) return rc;
    if (DL.nnz > INT_MAX) {
        free_device_layout(DL);
        return fail(SWE_EINVAL, "%lld entries do not fit the solver's 32-bit CSR", (long long)DL.nnz);
    }
    const int* lists[3] = {h_dir, ux_dir, uy_dir};
    const int nlist[3] = {n_h_dir, n_ux_dir, n_uy_dir};
    rc = pool_add_assembled(b->P, a, DL, bath, lists, nlist, *prm, un, t0, tfinal, max_steps, id, nullptr);
    free_device_layout(DL);
    return rc;
}

int swe_batch_set_bc(swe_batch* b, int id, int kind, double amp, double period, double level) {
    if (!b || !live(b->P, id)) return fail(SWE_EINVAL, "no instance %d", id);
    return pool_set_bc(b->P, id, kind, amp, period, level);
}

int swe_batch_set_bc_values(swe_batch* b, int id, const double* hv, const double* uxv, const double* uyv) {
    if (!b || !live(b->P, id)) return fail(SWE_EINVAL, "no instance %d", id);
    return set_bc_values(b->P, id, hv, uxv, uyv);
}

int swe_batch_remove(swe_batch* b, int id) {
    if (!b || !live(b->P, id)) return fail(SWE_EINVAL, "no instance %d", id);
    return pool_remove(b->P, id);
}

int swe_batch_step(swe_batch* b, int nsteps) {
    if (!b) return fail(SWE_EINVAL, "null batch");
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    return pool_queue_steps(b->P, nsteps);
}

int swe_batch_step_async(swe_batch* b, int nsteps, long long* ticket) {
    if (!b || !ticket) return fail(SWE_EINVAL, "null argument");
    if (nsteps < 0) return fail(SWE_EINVAL, "nsteps < 0");
    int rc = pool_queue_steps(b->P, nsteps);
    if (rc) return rc;
    return pool_snapshot(b->P, ticket);
}

int swe_batch_finished(swe_batch* b, long long ticket, int* ids, int cap, int* n) {
    if (!b || !n || (cap > 0 && !ids)) return fail(SWE_EINVAL, "null argument");
    return pool_finished(b->P, ticket, ids, cap, n);
}

int swe_batch_result(swe_batch* b, int id, double* un, int* state, double* time, long long* steps, int* error) {
    if (!b || !live(b->P, id)) return fail(SWE_EINVAL, "no instance %d", id);
    const InstHost& I = b->P->inst[(size_t)id];
    if (!I.final_known) return fail(SWE_EINVAL, "instance %d has not been reported finished", id);
    if (un) {
        int rc = pool_result(b->P, id, un);
        if (rc) return rc;
    }
    const InstDev& D = I.final_rec;
    const int f = I.final_flags;
    int st = D.state;
    if (st == INST_RUNNING && f) st = INST_FAILED;
    else if (st == INST_RUNNING) st = INST_DONE;
    if (state) *state = st;
    if (time) *time = D.time;
    if (steps) *steps = D.steps;
    if (error) *error = (f & FLAG_STEP) ? SWE_EHIP : (f & FLAG_LAMBDA) ? SWE_ELAMBDA : (f & FLAG_NEG_H) ? SWE_ENEG_H : SWE_OK;
    return SWE_OK;
}

int swe_batch_sync(swe_batch* b) {
    if (!b) return fail(SWE_EINVAL, "null batch");
    SWE_HIPCHK(hipStreamSynchronize(b->P->stream));
    return SWE_OK;
}

int swe_batch_status(swe_batch* b, int id, int* state, double* time, long long* steps, int* error) {
    if (!b || !live(b->P, id)) return fail(SWE_EINVAL, "no instance %d", id);
    int rc = pool_read_table(b->P);
    if (rc) return rc;
    const InstDev& D = b->P->tab_host[id];
    const int f = b->P->flags_host[id];
    int st = D.state;
    // an instance whose stopping rule holds is done even before the next step's bookkeeping
    if (st == INST_RUNNING && f) st = INST_FAILED;
    else if (st == INST_RUNNING && !(D.time < D.tfinal && D.steps < D.max_steps)) st = INST_DONE;
    if (state) *state = st;
    if (time) *time = D.time;
    if (steps) *steps = D.steps;
    if (error) *error = (f & FLAG_STEP) ? SWE_EHIP : (f & FLAG_LAMBDA) ? SWE_ELAMBDA : (f & FLAG_NEG_H) ? SWE_ENEG_H : SWE_OK;
    return SWE_OK;
}

int swe_batch_poll(swe_batch* b, int* ids, int cap, int* n) {
    if (!b || !n || (cap > 0 && !ids)) return fail(SWE_EINVAL, "null argument");
    *n = 0;
    int rc = pool_read_table(b->P);
    if (rc) return rc;
    Pool* P = b->P;
    const int ni = P->ninst();
    bool unlisted = false;
    for (int k = 0; k < ni; ++k) {
        if (!P->inst[(size_t)k].used) continue;
        const InstDev& D = P->tab_host[k];
        const bool over = D.state != INST_RUNNING || P->flags_host[k] != 0 ||
                          !(D.time < D.tfinal && D.steps < D.max_steps);
        if (!over) continue;
        if (P->listed[(size_t)k]) {          // it takes no more steps: its waves leave the launches
            P->listed[(size_t)k] = 0;
            unlisted = true;
        }
        if (*n < cap) ids[(*n)++] = k;
    }
    return unlisted ? pool_update_active(P) : SWE_OK;
}

int swe_batch_errors_create(swe_batch* b, int id, int me, const int* jj, const double* rr, swe_errors** out) {
    clear_error();
    if (!b || !jj || !rr || !out) return fail(SWE_EINVAL, "null argument");
    *out = nullptr;
    if (!live(b->P, id)) return fail(SWE_EINVAL, "no instance %d", id);
    swe::ErrEval* E = nullptr;
    const int rc = errors_create(b->P, id, me, jj, rr, &E);
    if (rc) return rc;
    *out = new swe_errors{nullptr, b->P, id, b->P->inst[(size_t)id].since, E};
    return SWE_OK;
}

int swe_batch_get_state(swe_batch* b, int id, double* un) {
    if (!b || !un || !live(b->P, id)) return fail(SWE_EINVAL, "no instance %d", id);
    return pool_download_state(b->P, id, b->P->S[0], un);
}

int swe_batch_free(swe_batch* b, long long* rows, long long* entries, int* instances) {
    if (!b) return fail(SWE_EINVAL, "null batch");
    if (rows) *rows = b->P->a_slices.largest_free() * 64;
    if (entries) *entries = b->P->a_slots.largest_free();
    if (instances) {
        int n = 0;
        for (const auto& q : b->P->inst) n += q.used ? 0 : 1;
        *instances = n;
    }
    return SWE_OK;
}

int swe_batch_tune(swe_batch* b, int steps) {
    if (!b) return fail(SWE_EINVAL, "null batch");
    return pool_tune(b->P, steps, 2);
}

int swe_batch_use_graph(swe_batch* b, int on) {
    if (!b) return fail(SWE_EINVAL, "null batch");
    b->P->use_graph = on != 0;
    return SWE_OK;
}

void* swe_batch_stream(swe_batch* b) { return b ? (void*)b->P->stream : nullptr; }

int swe_batch_enable_kernel_timing(swe_batch* b, const char* kernel) {
    if (!b) return fail(SWE_EINVAL, "null batch");
    return timing_enable(b->P, kernel);
}

int swe_batch_kernel_timing(swe_batch* b, const char* kernel, int* launches, double* ms) {
    if (!b || !kernel || !launches || !ms) return fail(SWE_EINVAL, "null argument");
    return timing_read(b->P, kernel, launches, ms);
}

}  // extern "C"
