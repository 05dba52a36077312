// swe_engine.cpp -- the pool: an HBM arena of meshes advanced together by the step kernels.
//
// Admission builds a mesh's sliced layout on the host, takes first-fit ranges of slices, entry
// slots, compact off-diagonal lists and boundary entries, uploads the bulk data on a copy stream
// and then, on the compute stream, writes the instance's table entry, which is what makes the
// step kernels see it.  Removal clears the slice and boundary owners on the compute stream and
// records an event that the next upload waits for, so a range is never rewritten while queued
// kernels of its previous owner may still run.  A step is a fixed launch sequence over the list
// of listed slices (every admitted instance until it is found done or failed); an instance that
// stops between list updates skips the step per wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "swe_engine.h"

namespace swe {

namespace {
thread_local std::string g_err;
}  // namespace

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
const char* last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

// ---- RangeAlloc -----------------------------------------------------------------------------
void RangeAlloc::init(int64_t c) {
    free_.clear();
    cap = c;
    if (c > 0) free_[0] = c;
}

int64_t RangeAlloc::alloc(int64_t n) {
    if (n <= 0) return 0;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= n) {
            const int64_t s = it->first, l = it->second;
            free_.erase(it);
            if (l > n) free_[s + n] = l - n;
            return s;
        }
    }
    return -1;
}

void RangeAlloc::release(int64_t start, int64_t n) {
    if (n <= 0) return;
    auto it = free_.emplace(start, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

int64_t RangeAlloc::high_water() const {
    if (free_.empty()) return cap;
    auto last = std::prev(free_.end());
    return (last->first + last->second == cap) ? last->first : cap;
}

int64_t RangeAlloc::largest_free() const {
    int64_t m = 0;
    for (const auto& kv : free_) m = std::max(m, kv.second);
    return m;
}

// ---- parameters -----------------------------------------------------------------------------
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
    // compute_lambda_vacc knows tests 1 to 13 and STOPs on any other (SOURCES/condlim.f90:897,952-954)
    if (p.type_test < 1 || p.type_test > 13)
        return fail(SWE_EBADOPT, "BUG in compute_lambda: type_test %d is not one of 1..13", p.type_test);
    // test 13 is the 5-component SGN system (SOURCES/main.f90:24-26): with entropy viscosity or FCT
    // the reference's entropy_residual reads Entprime(4:5,:), which it never sets
    // (SOURCES/update.f90:853-862), so only the galerkin and first-order schemes are defined for it
    if (p.type_test == 13 && (p.viscosity_type == SWE_ENTROPY_VISC || p.viscosity_type == SWE_FCT))
        return fail(SWE_EBADOPT, "type_test 13 (SGN model): entropy_residual reads Entprime(4:5,:), which the "
                                 "reference never sets; use galerkin or viscous");
    K = Consts{};
    K.g = p.gravity;
    // test 9: the driver sets htiny = 1e-3 and dt = 2.5e-2 at the top of every step instead of
    // calling COMPUTE_DT (SOURCES/main.f90:67-69); the library uses that htiny in every call of this
    // test and that dt in its own time loop (the only reference call with another htiny is the
    // COMPUTE_DT before the loop, main.f90:40, whose dt the loop replaces)
    K.htiny = (p.type_test == 9) ? 1.0e-3 : p.htiny;
    K.dt_fixed = (p.type_test == 9) ? 2.5e-2 : 0.0;
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
    K.check_h = 1;
    K.lambda_sgn = p.lambda_sgn;
    K.src = (p.type_test == 9 || p.type_test == 10) ? SRC_MANNING
            : (p.type_test == 11)                   ? SRC_LINEAR
            : (p.type_test == 12)                   ? SRC_CORIOLIS
            : (p.type_test == 13)                   ? SRC_SGN
                                                    : SRC_NONE;
    return SWE_OK;
}

int mode_of(const swe_params& p, Mode& m) {
    Consts K;
    const int rc = make_consts(p, K);
    if (rc) return rc;
    m.viscosity_type = p.viscosity_type;
    m.viscous_type = p.viscous_type;
    m.alpha = p.alpha_limit ? 1 : 0;
    m.roundoff = p.roundoff_fix ? 1 : 0;
    m.sgn = (p.type_test == 13) ? 1 : 0;
    return SWE_OK;
}

bool same_mode(const Mode& a, const Mode& b) {
    return a.viscosity_type == b.viscosity_type && a.viscous_type == b.viscous_type && a.alpha == b.alpha &&
           a.roundoff == b.roundoff && a.sgn == b.sgn;
}

// ---- host layout ----------------------------------------------------------------------------
namespace {

// f(begin, end) over [0, n) in contiguous chunks on the host's cores (serially for small n).
template <class F>
void parallel_for(int64_t n, F&& f, int64_t serial_below = 1 << 15) {
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < serial_below || nt == 1) {
        if (n > 0) f((int64_t)0, n);
        return;
    }
    const int64_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        const int64_t b = (int64_t)t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([=, &f] { f(b, e); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

int validate_mesh(const swe_mesh_desc& d) {
    if (d.np <= 0 || d.nnz <= 0 || !d.ia || !d.ja || !d.cx || !d.cy || !d.lumped || !d.bath)
        return fail(SWE_EINVAL, "mesh descriptor is incomplete");
    const int np = d.np, nnz = d.nnz;
    if (d.ia[0] != 1 || d.ia[np] - 1 != nnz) return fail(SWE_EINVAL, "ia must start at 1 and end at nnz+1");
    for (int i = 0; i < np; ++i)
        if (d.ia[i + 1] < d.ia[i]) return fail(SWE_EINVAL, "ia is not monotone at row %d", i + 1);
    // rows checked in parallel; the first bad row (in row order) is reported
    auto row_ok = [&](int i) {
        bool has_diag = false;
        for (int p = d.ia[i] - 1; p < d.ia[i + 1] - 1; ++p) {
            const int j = d.ja[p];
            if (j < 1 || j > np) return false;
            has_diag |= (j - 1 == i);
        }
        return has_diag && d.lumped[i] > 0.0;
    };
    std::atomic<int> bad{np};
    parallel_for(np, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i)
            if (!row_ok((int)i)) {
                int cur = bad.load();
                while (i < cur && !bad.compare_exchange_weak(cur, (int)i)) {
                }
                return;
            }
    });
    if (bad.load() < np) {
        const int i = bad.load();
        for (int p = d.ia[i] - 1; p < d.ia[i + 1] - 1; ++p)
            if (d.ja[p] < 1 || d.ja[p] > np) return fail(SWE_EINVAL, "ja(%d)=%d out of range", p + 1, d.ja[p]);
        if (!(d.lumped[i] > 0.0)) {
            bool has_diag = false;
            for (int p = d.ia[i] - 1; p < d.ia[i + 1] - 1; ++p) has_diag |= (d.ja[p] - 1 == i);
            if (has_diag) return fail(SWE_EINVAL, "lumped(%d) must be > 0", i + 1);
        }
        return fail(SWE_EINVAL, "row %d has no diagonal entry", i + 1);
    }
    const int n[3] = {d.n_h_dir, d.n_ux_dir, d.n_uy_dir};
    const int* l[3] = {d.h_dir, d.ux_dir, d.uy_dir};
    const char* names[3] = {"h_js_D", "ux_js_D", "uy_js_D"};
    for (int k = 0; k < 3; ++k) {
        if (n[k] < 0 || (n[k] > 0 && !l[k])) return fail(SWE_EINVAL, "%s list is malformed", names[k]);
        for (int q = 0; q < n[k]; ++q)
            if (l[k][q] < 1 || l[k][q] > np) return fail(SWE_EINVAL, "%s node %d out of range", names[k], l[k][q]);
    }
    return SWE_OK;
}

namespace {

// Mean |i - j| over the off-diagonal entries of the CSR graph in the numbering `pos`
// (pos[i] = position of row i; identity when null).
double mean_offset(const swe_mesh_desc& d, const std::vector<int>* pos) {
    std::atomic<int64_t> sum{0}, n{0};
    parallel_for(d.np, [&](int64_t b0, int64_t e0) {
        int64_t s = 0, c = 0;
        for (int64_t i = b0; i < e0; ++i)
            for (int p = d.ia[i] - 1; p < d.ia[i + 1] - 1; ++p) {
                const int j = d.ja[p] - 1;
                if (j == (int)i) continue;
                const int a = pos ? (*pos)[(size_t)i] : (int)i, b = pos ? (*pos)[(size_t)j] : j;
                s += (a > b) ? a - b : b - a;
                ++c;
            }
        sum += s;
        n += c;
    });
    return n.load() ? (double)sum.load() / (double)n.load() : 0.0;
}

// Reverse Cuthill-McKee order of the CSR graph (each component from a pseudo-peripheral node, the
// end of a breadth-first sweep from a node of least degree; neighbours by increasing degree).
std::vector<int> rcm_order(const swe_mesh_desc& d) {
    const int np = d.np;
    auto deg = [&](int i) { return d.ia[i + 1] - d.ia[i]; };
    std::vector<int> order;
    order.reserve((size_t)np);
    std::vector<char> seen((size_t)np, 0);
    std::vector<int> level((size_t)np, -1), nb;
    auto bfs = [&](int s, bool record) {
        // breadth-first sweep from s; returns the last node reached
        std::vector<int> q{s};
        level[(size_t)s] = 0;
        int last = s;
        for (size_t h = 0; h < q.size(); ++h) {
            const int u = q[h];
            last = u;
            nb.clear();
            for (int p = d.ia[u] - 1; p < d.ia[u + 1] - 1; ++p) {
                const int v = d.ja[p] - 1;
                if (v != u && level[(size_t)v] < 0) {
                    level[(size_t)v] = level[(size_t)u] + 1;
                    nb.push_back(v);
                }
            }
            std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg(a) < deg(b); });
            q.insert(q.end(), nb.begin(), nb.end());
        }
        if (record)
            for (int u : q) {
                seen[(size_t)u] = 1;
                order.push_back(u);
            }
        else
            for (int u : q) level[(size_t)u] = -1;
        return last;
    };
    std::vector<int> by_deg((size_t)np);
    for (int i = 0; i < np; ++i) by_deg[(size_t)i] = i;
    std::stable_sort(by_deg.begin(), by_deg.end(), [&](int a, int b) { return deg(a) < deg(b); });
    for (int s : by_deg) {
        if (seen[(size_t)s]) continue;
        const int far = bfs(s, false);
        bfs(far, true);
    }
    std::reverse(order.begin(), order.end());
    return order;
}

// SWE_REORDER: 0 keeps the caller's order, 2 always renumbers, otherwise (default) the rows are
// renumbered by reverse Cuthill-McKee when the caller's numbering scatters neighbours (mean offset
// above 2*sqrt(np): a lattice numbered row by row has about 0.7*sqrt(np), the loader's first-touch
// numbering of an unstructured mesh about 3.4*sqrt(np)) and the new order is better.
}  // namespace

int reorder_mode() {
    const char* v = std::getenv("SWE_REORDER");
    const int m = v ? std::atoi(v) : 1;
    return (m >= 0 && m <= 2) ? m : 1;
}

// Internal numbering: a base order of the rows (the caller's, or reverse Cuthill-McKee when the
// caller's numbering scatters neighbours), then inside consecutive windows of `sort_window` rows
// of it, rows stably sorted by length, so the 64 rows of a slice (one wave) mostly have equal
// length.  Only the processing order and placement change; each row keeps its entries in the
// caller's order, which is the order of every row sum.
int build_layout(const swe_mesh_desc& d, int sort_window, HostLayout& L) {
    int rc = validate_mesh(d);
    if (rc) return rc;
    const int np = d.np, nnz = d.nnz;
    L = HostLayout{};
    L.np = np;
    L.nnz = nnz;
    const int* ia = d.ia;
    auto rowlen = [&](int i) { return ia[i + 1] - ia[i]; };
    L.perm.resize((size_t)np);
    L.iperm.resize((size_t)np);
    std::vector<int> base((size_t)np);
    for (int i = 0; i < np; ++i) base[(size_t)i] = i;
    const int mode = reorder_mode();
    if (mode == 2 || (mode == 1 && mean_offset(d, nullptr) > 2.0 * std::sqrt((double)np))) {
        std::vector<int> rcm = rcm_order(d);
        std::vector<int> pos((size_t)np);
        for (int r = 0; r < np; ++r) pos[(size_t)rcm[(size_t)r]] = r;
        if (mode == 2 || mean_offset(d, &pos) < mean_offset(d, nullptr)) base.swap(rcm);
    }
    const int W = std::max(1, sort_window);
    const int nwin = (np + W - 1) / W;
    parallel_for(nwin, [&](int64_t b, int64_t e) {
        for (int64_t w = b; w < e; ++w) {
            const int w0 = (int)w * W, w1 = std::min(np, w0 + W);
            for (int i = w0; i < w1; ++i) L.perm[(size_t)i] = base[(size_t)i];
            std::stable_sort(L.perm.begin() + w0, L.perm.begin() + w1,
                             [&](int a, int c) { return rowlen(a) < rowlen(c); });
        }
    }, 512);
    L.np_pad = (np + 63) / 64 * 64;
    L.nsl = L.np_pad / 64;
    L.len.assign((size_t)L.np_pad, 0);
    L.sptr.assign((size_t)L.nsl + 1, 0);
    parallel_for(np, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; ++r) {
            L.iperm[(size_t)L.perm[(size_t)r]] = (int)r;
            L.len[(size_t)r] = rowlen(L.perm[(size_t)r]);
        }
    });
    for (int sl = 0; sl < L.nsl; ++sl) {
        int w = 0;
        for (int l = 0; l < 64; ++l) w = std::max(w, L.len[(size_t)sl * 64 + l]);
        L.sptr[(size_t)sl + 1] = L.sptr[(size_t)sl] + 64 * (int64_t)w;
        L.wmax = std::max(L.wmax, w);
    }
    if (L.wmax > kMaxRow) return fail(SWE_EINVAL, "a row has %d entries; at most %d are supported", L.wmax, kMaxRow);
    L.nslot = L.sptr[(size_t)L.nsl];
    L.col.assign((size_t)L.nslot, 0);
    L.dcol.assign((size_t)L.nslot, 0);
    L.c.assign((size_t)L.nslot, make_double2(0.0, 0.0));
    L.csr2slot.assign((size_t)nnz, 0);
    L.m.assign((size_t)np, 0.0);
    L.z.assign((size_t)np, 0.0);
    L.dpos.assign((size_t)L.np_pad, 0);
    parallel_for(np, [&](int64_t rb, int64_t re) {
        for (int r = (int)rb; r < (int)re; ++r) {
            const int i = L.perm[(size_t)r];
            const int64_t base = L.sptr[(size_t)(r >> 6)] + (r & 63);
            const int p0 = ia[i] - 1;
            for (int q = 0; q < L.len[(size_t)r]; ++q) {
                const int p = p0 + q;
                const int j = d.ja[p] - 1;
                if (j == i) L.dpos[(size_t)r] = q;
                const int64_t slot = base + 64 * (int64_t)q;
                L.col[(size_t)slot] = L.iperm[(size_t)j];
                const int off = L.iperm[(size_t)j] - r;
                L.dcol[(size_t)slot] = (int16_t)((off > kColFar && off <= 32767) ? off : kColFar);
                L.c[(size_t)slot] = make_double2(d.cx[p], d.cy[p]);
                L.csr2slot[(size_t)p] = slot;
            }
            L.m[(size_t)r] = d.lumped[i];
            L.z[(size_t)r] = d.bath[i];
        }
    });
    // compact per-slice lists of off-diagonal slots, in slot order (coalesced in the dt pass):
    // counted per slice, offsets by a prefix sum, then filled per slice
    L.eoff.assign((size_t)L.nsl + 1, 0);
    auto slice_entries = [&](int sl, uint16_t* out) {
        const int w = (int)((L.sptr[(size_t)sl + 1] - L.sptr[(size_t)sl]) / 64);
        int64_t n = 0;
        for (int q = 0; q < w; ++q)
            for (int l = 0; l < 64; ++l) {
                const int r = sl * 64 + l;
                if (r < np && q < L.len[(size_t)r] && q != L.dpos[(size_t)r]) {
                    if (out) out[n] = (uint16_t)(q * 64 + l);
                    ++n;
                }
            }
        return n;
    };
    parallel_for(L.nsl, [&](int64_t b, int64_t e) {
        for (int64_t sl = b; sl < e; ++sl) L.eoff[(size_t)sl + 1] = slice_entries((int)sl, nullptr);
    }, 512);
    for (int sl = 0; sl < L.nsl; ++sl) L.eoff[(size_t)sl + 1] += L.eoff[(size_t)sl];
    L.elist.assign((size_t)L.eoff[(size_t)L.nsl], 0);
    parallel_for(L.nsl, [&](int64_t b, int64_t e) {
        for (int64_t sl = b; sl < e; ++sl) slice_entries((int)sl, L.elist.data() + L.eoff[(size_t)sl]);
    }, 512);
    const int* src[3] = {d.h_dir, d.ux_dir, d.uy_dir};
    const int ns[3] = {d.n_h_dir, d.n_ux_dir, d.n_uy_dir};
    for (int k = 0; k < 3; ++k) {
        L.lists[k].resize((size_t)ns[k]);
        for (int q = 0; q < ns[k]; ++q) L.lists[k][(size_t)q] = L.iperm[(size_t)src[k][q] - 1];
    }
    return SWE_OK;
}

// ---- pool -----------------------------------------------------------------------------------
namespace {

template <class T>
int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    const hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(SWE_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n * sizeof(T), hipGetErrorString(e));
    return SWE_OK;
}

// Device bytes per unit of each capacity.
constexpr size_t kRowBytes = sizeof(int) + 2 * sizeof(uint8_t) + sizeof(double) * 2 + sizeof(Node) * 4 + sizeof(double) * 3 +
                             sizeof(double) * 4 + sizeof(double) * 3;   // perm, len dpos, m z, S[3] scratch, ulo, alpha res Rm nwork, stage3
constexpr size_t kSlotBytes = sizeof(int) + sizeof(int16_t) + sizeof(double2) * 2 + sizeof(double) * 2;  // col dcol c dmu d0 ework
constexpr size_t kSliceBytes = sizeof(SliceInfo);
constexpr size_t kInstBytes = sizeof(InstDev) + sizeof(unsigned long long) + sizeof(int);
constexpr size_t kBcBytes = sizeof(BcEntry) + 3 * sizeof(double);   // entry, value of each stage

void free_pool(Pool* P) {
    if (!P) return;
    void* bufs[] = {P->sl,  P->len,     P->dpos,  P->perm, P->m,     P->z,     P->col,   P->dcol,  P->c,
                    P->elist,   P->S[0],  P->S[1], P->S[2],  P->scratch, P->ulo_h, P->ulo_q, P->d0,    P->dmu,
                    P->ework,   P->alpha, P->res,  P->Rm,    P->nwork, P->stage3, P->tab, P->vmax,
                    P->flags, P->bc,    P->bcv,  P->X[0], P->X[1], P->X[2], P->X[3], P->stage5, P->xscratch,
                    P->bar};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (P->act) (void)hipFree(P->act);
    for (int k = 0; k < 2; ++k) {
        if (P->act_host[k]) (void)hipHostFree(P->act_host[k]);
        if (P->act_ev[k]) (void)hipEventDestroy(P->act_ev[k]);
    }
    for (auto& sn : P->snaps) {
        if (sn.tab) (void)hipHostFree(sn.tab);
        if (sn.flags) (void)hipHostFree(sn.flags);
        if (sn.ev) (void)hipEventDestroy(sn.ev);
    }
    if (P->tab_host) (void)hipHostFree(P->tab_host);
    if (P->flags_host) (void)hipHostFree(P->flags_host);
    if (P->gexec) (void)hipGraphExecDestroy(P->gexec);
    for (hipEvent_t e : P->tk.ev) (void)hipEventDestroy(e);
    if (P->released) (void)hipEventDestroy(P->released);
    if (P->copy) (void)hipStreamDestroy(P->copy);
    if (P->stream) (void)hipStreamDestroy(P->stream);
    delete P;
}

}  // namespace

int Pool::ninst() const {
    for (int k = (int)inst.size() - 1; k >= 0; --k)
        if (inst[(size_t)k].used) return k + 1;
    return 0;
}

size_t pool_bytes(const PoolCaps& c) {
    const int64_t slices = (c.rows + 63) / 64;
    return (size_t)slices * 64 * kRowBytes + (size_t)c.slots * kSlotBytes + (size_t)slices * kSliceBytes +
           (size_t)c.elist * sizeof(uint16_t) + (size_t)c.instances * kInstBytes + (size_t)c.bc * kBcBytes;
}

int sort_window_default() {
    const char* v = std::getenv("SWE_SORT_WINDOW");
    const int w = v ? std::atoi(v) : 0;
    return (w >= 1 && w <= (1 << 20)) ? w : 64;
}

int pool_create(const PoolCaps& caps, Pool** out) {
    *out = nullptr;
    if (caps.rows <= 0 || caps.slots <= 0 || caps.instances <= 0 || caps.elist < 0 || caps.bc < 0)
        return fail(SWE_EINVAL, "pool capacities must be positive");
    if (caps.rows > (int64_t)INT_MAX - 64) return fail(SWE_EINVAL, "at most 2^31 rows per pool");
    Pool* P = new Pool();
    P->sort_window = sort_window_default();
    const int64_t slices = (caps.rows + 63) / 64;
    P->rows_cap = slices * 64;
    P->slots_cap = caps.slots;
    P->elist_cap = caps.elist;
    P->bc_cap = caps.bc;
    P->inst_cap = caps.instances;
    P->a_slices.init(slices);
    P->a_slots.init(caps.slots);
    P->a_elist.init(caps.elist);
    P->a_bc.init(caps.bc);
    P->inst.assign((size_t)caps.instances, InstHost{});
    hipError_t e = hipGetDevice(&P->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&P->copy, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&P->released, hipEventDisableTiming);
    if (e != hipSuccess) {
        free_pool(P);
        return fail(SWE_EHIP, "pool streams: %s", hipGetErrorString(e));
    }
    const size_t R = (size_t)P->rows_cap, S = (size_t)caps.slots;
    int rc;
#define ALLOC(ptr, n)                      \
    do {                                   \
        if ((rc = dalloc(&(ptr), (n)))) {  \
            free_pool(P);                  \
            return rc;                     \
        }                                  \
    } while (0)
    ALLOC(P->sl, (size_t)slices);
    ALLOC(P->len, R);
    ALLOC(P->dpos, R);
    ALLOC(P->perm, R);
    ALLOC(P->m, R);
    ALLOC(P->z, R);
    ALLOC(P->col, S);
    ALLOC(P->dcol, S);
    ALLOC(P->c, S);
    ALLOC(P->elist, (size_t)std::max<int64_t>(caps.elist, 1));
    for (int k = 0; k < 3; ++k) ALLOC(P->S[k], R);
    ALLOC(P->scratch, R);
    ALLOC(P->ulo_h, R);
    ALLOC(P->ulo_q, R);
    ALLOC(P->d0, S);
    ALLOC(P->dmu, S);
    ALLOC(P->ework, S);
    ALLOC(P->alpha, R);
    ALLOC(P->res, R);
    ALLOC(P->Rm, R);
    ALLOC(P->nwork, R);
    ALLOC(P->stage3, 3 * R);
    ALLOC(P->tab, (size_t)caps.instances);
    ALLOC(P->vmax, (size_t)caps.instances);
    ALLOC(P->flags, (size_t)caps.instances);
    ALLOC(P->bc, (size_t)std::max<int64_t>(caps.bc, 1));
    P->bcv_stride = std::max<int64_t>(caps.bc, 1);
    ALLOC(P->bcv, 3 * (size_t)P->bcv_stride);
    ALLOC(P->act, (size_t)slices);
    ALLOC(P->bar, 2);
#undef ALLOC
    P->listed.assign((size_t)caps.instances, 0);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipHostMalloc((void**)&P->act_host[k], sizeof(int) * (size_t)slices, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&P->act_ev[k], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(P->act_ev[k], P->stream);
    }
    if (e != hipSuccess) {
        free_pool(P);
        return fail(SWE_EHIP, "pool lists: %s", hipGetErrorString(e));
    }
    for (auto& sn : P->snaps) {
        if (e == hipSuccess) e = hipHostMalloc((void**)&sn.tab, sizeof(InstDev) * (size_t)caps.instances, hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sn.flags, sizeof(int) * (size_t)caps.instances, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sn.ev, hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipHostMalloc((void**)&P->tab_host, sizeof(InstDev) * (size_t)caps.instances, hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&P->flags_host, sizeof(int) * (size_t)caps.instances, hipHostMallocDefault);
    if (e != hipSuccess) {
        free_pool(P);
        return fail(SWE_ENOMEM, "hipHostMalloc failed: %s", hipGetErrorString(e));
    }
    // empty tables: every slice free, every instance free; len = 0 marks rows without work
    std::vector<SliceInfo> sl0((size_t)slices);
    for (auto& s : sl0) s = SliceInfo{0, 0, -1, 0, {0, 0}};
    std::vector<InstDev> t0((size_t)caps.instances);
    std::memset(t0.data(), 0, sizeof(InstDev) * t0.size());
    e = hipMemcpy(P->sl, sl0.data(), sizeof(SliceInfo) * sl0.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(P->tab, t0.data(), sizeof(InstDev) * t0.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(P->len, 0, sizeof(uint8_t) * R);
    if (e == hipSuccess) e = hipMemset(P->vmax, 0, sizeof(unsigned long long) * (size_t)caps.instances);
    if (e == hipSuccess) e = hipMemset(P->flags, 0, sizeof(int) * (size_t)caps.instances);
    if (e == hipSuccess) e = hipMemset(P->bar, 0, 2 * sizeof(unsigned));
    int coop = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, P->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&P->num_cu, hipDeviceAttributeMultiprocessorCount, P->device);
    P->coop_ok = coop != 0;
    if (e == hipSuccess) e = hipMemset(P->bc, 0xff, sizeof(BcEntry) * (size_t)std::max<int64_t>(caps.bc, 1));
    if (e == hipSuccess) e = hipEventRecord(P->released, P->stream);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_pool(P);
        return fail(SWE_EHIP, "pool init: %s", hipGetErrorString(e));
    }
    *out = P;
    return SWE_OK;
}

void pool_destroy(Pool* P) {
    if (!P) return;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    if (P->copy) (void)hipStreamSynchronize(P->copy);
    free_pool(P);
}

NodeX* ext_of(const Pool& P, const Node* s) {
    for (int q = 0; q < 3; ++q)
        if (s == P.S[q]) return P.X[q];
    return (s == P.scratch) ? P.X[3] : nullptr;
}

// The SGN model's h*eta and h*w arrays and its staging buffer, made when the first SGN instance is
// admitted (pools that never hold one do without them).
int ensure_sgn_buffers(Pool* P) {
    if (P->stage5) return SWE_OK;
    const size_t R = (size_t)P->rows_cap;
    int rc;
    for (int q = 0; q < 4; ++q)
        if ((rc = dalloc(&P->X[q], R))) return rc;
    if ((rc = dalloc(&P->xscratch, R))) return rc;
    if ((rc = dalloc(&P->stage5, 5 * R))) return rc;
    for (int q = 0; q < 4; ++q) SWE_HIPCHK(hipMemset(P->X[q], 0, sizeof(NodeX) * R));
    return SWE_OK;
}

// bdy of the SGN model writes h*eta and h*w on the ux and uy lists with sol_anal(4) and (5)
// (SOURCES/main.f90:184-187), which the reference defines only at t <= 1e-10
// (SOURCES/condlim.f90:801-826): SGN instances take h lists only.
int check_sgn_lists(const Mode& md, int nux, int nuy) {
    if (md.sgn && (nux > 0 || nuy > 0))
        return fail(SWE_EBADOPT, "type_test 13 (SGN model): bdy of h*eta and h*w on the ux/uy lists uses sol_anal(4:5), "
                                 "undefined after t = 1e-10 in the reference; give h Dirichlet nodes only");
    return SWE_OK;
}

int pool_add(Pool* P, const swe_mesh_desc& d, const swe_params& prm, const double* un, double t0, double tfinal,
             long long max_steps, int* id, HostLayout* keep) {
    Mode md;
    int rc = mode_of(prm, md);
    if (rc) return rc;
    if (P->mode_set && !same_mode(md, P->mode))
        return fail(SWE_EBADOPT, "the pool runs viscosity %d/type%d/alpha %d/roundoff %d; this instance differs",
                    P->mode.viscosity_type, P->mode.viscous_type, P->mode.alpha, P->mode.roundoff);
    Consts K;
    if ((rc = make_consts(prm, K))) return rc;
    if ((rc = check_sgn_lists(md, d.n_ux_dir, d.n_uy_dir))) return rc;
    if (md.sgn && (rc = ensure_sgn_buffers(P))) return rc;
    HostLayout Lbuf;
    HostLayout& L = keep ? *keep : Lbuf;
    if ((rc = build_layout(d, P->sort_window, L))) return rc;
    int k = -1;
    for (int q = 0; q < P->inst_cap; ++q)
        if (!P->inst[(size_t)q].used) { k = q; break; }
    const int64_t nbc = (int64_t)L.lists[0].size() + (int64_t)L.lists[1].size() + (int64_t)L.lists[2].size();
    const int64_t ne = (int64_t)L.elist.size();
    const int64_t s0 = (k < 0) ? -1 : P->a_slices.alloc(L.nsl);
    const int64_t p0 = (s0 < 0) ? -1 : P->a_slots.alloc(L.nslot);
    const int64_t e0 = (p0 < 0) ? -1 : P->a_elist.alloc(ne);
    const int64_t b0 = (e0 < 0) ? -1 : P->a_bc.alloc(nbc);
    if (k < 0 || s0 < 0 || p0 < 0 || e0 < 0 || b0 < 0) {
        if (s0 >= 0) P->a_slices.release(s0, L.nsl);
        if (p0 >= 0) P->a_slots.release(p0, L.nslot);
        if (e0 >= 0) P->a_elist.release(e0, ne);
        return fail(SWE_EFULL, "the pool has no room for a mesh of %d nodes now", d.np);
    }
    const int64_t r0 = s0 * 64;
    // global arrays of this instance
    std::vector<SliceInfo> sls((size_t)L.nsl);
    for (int s = 0; s < L.nsl; ++s)
        sls[(size_t)s] = SliceInfo{p0 + L.sptr[(size_t)s], e0 + L.eoff[(size_t)s], k,
                                   (int)(L.eoff[(size_t)s + 1] - L.eoff[(size_t)s]), {0, 0}};
    std::vector<uint8_t> len8(L.len.begin(), L.len.end()), dpos8(L.dpos.begin(), L.dpos.end());
    std::vector<int> colg((size_t)L.nslot);
    for (int64_t q = 0; q < L.nslot; ++q) colg[(size_t)q] = L.col[(size_t)q] + (int)r0;
    std::vector<BcEntry> bce((size_t)nbc);
    {
        int64_t q = 0;
        for (int comp = 0; comp < 3; ++comp)
            for (int node : L.lists[comp]) {
                bce[(size_t)q] = BcEntry{node + (int)r0, k, comp, (int)(b0 + q)};
                ++q;
            }
    }
    // bulk upload on the copy stream, after any release of these ranges has taken effect
    hipStream_t cs = P->copy;
    auto up = [&](void* dst, const void* src, size_t bytes) -> hipError_t {
        if (bytes == 0) return hipSuccess;
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, cs);
    };
    hipError_t e = hipStreamWaitEvent(cs, P->released, 0);
    if (e == hipSuccess) e = up(P->len + r0, len8.data(), len8.size());
    if (e == hipSuccess) e = up(P->dpos + r0, dpos8.data(), dpos8.size());
    if (e == hipSuccess) e = up(P->perm + r0, L.perm.data(), sizeof(int) * L.perm.size());
    if (e == hipSuccess) e = up(P->m + r0, L.m.data(), sizeof(double) * L.m.size());
    if (e == hipSuccess) e = up(P->z + r0, L.z.data(), sizeof(double) * L.z.size());
    if (e == hipSuccess) e = up(P->col + p0, colg.data(), sizeof(int) * colg.size());
    if (e == hipSuccess) e = up(P->dcol + p0, L.dcol.data(), sizeof(int16_t) * L.dcol.size());
    if (e == hipSuccess) e = up(P->c + p0, L.c.data(), sizeof(double2) * L.c.size());
    if (e == hipSuccess) e = up(P->elist + e0, L.elist.data(), sizeof(uint16_t) * L.elist.size());
    if (e == hipSuccess) e = up(P->bc + b0, bce.data(), sizeof(BcEntry) * bce.size());
    if (e == hipSuccess) e = up(P->sl + s0, sls.data(), sizeof(SliceInfo) * sls.size());
    if (e == hipSuccess && un) {
        if (md.sgn) e = up(P->stage5 + 5 * r0, un, sizeof(double) * 5 * (size_t)d.np);
        else e = up(P->stage3 + 3 * r0, un, sizeof(double) * 3 * (size_t)d.np);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(cs);
    if (e != hipSuccess) {
        P->a_slices.release(s0, L.nsl);
        P->a_slots.release(p0, L.nslot);
        P->a_elist.release(e0, ne);
        P->a_bc.release(b0, nbc);
        return fail(SWE_EHIP, "upload of an admitted mesh: %s", hipGetErrorString(e));
    }
    InstHost I;
    I.np = d.np;
    I.nnz = d.nnz;
    I.np_pad = L.np_pad;
    I.wmax = L.wmax;
    I.slice0 = s0;
    I.nslices = L.nsl;
    I.slot0 = p0;
    I.nslots = L.nslot;
    I.e0 = e0;
    I.ne = ne;
    I.bc0 = b0;
    I.nbc = nbc;
    I.nh = (int)L.lists[0].size();
    I.nux = (int)L.lists[1].size();
    I.nuy = (int)L.lists[2].size();
    I.nc = md.sgn ? 5 : 3;
    if ((rc = pool_activate(P, k, I, prm, un != nullptr, t0, tfinal, max_steps))) return rc;
    *id = k;
    return SWE_OK;
}

int pool_activate(Pool* P, int k, const InstHost& layout, const swe_params& prm, bool packed_state, double t0,
                  double tfinal, long long max_steps) {
    Mode md;
    int rc = mode_of(prm, md);
    if (rc) return rc;
    Consts K;
    if ((rc = make_consts(prm, K))) return rc;
    InstHost& I = P->inst[(size_t)k];
    I = layout;
    I.used = true;
    I.final_known = false;
    I.since = P->next_ticket;
    I.prm = prm;
    const int64_t r0 = I.slice0 * 64;
    P->wmax = std::max(P->wmax, I.wmax);
    if (!P->mode_set) {
        P->mode = md;
        P->mode_set = true;
    }
    // the state, then the table entry: from the next queued step on the instance is live
    I.nc = md.sgn ? 5 : 3;
    if (packed_state) {
        if (md.sgn) launch_pack5(*P, I, P->stage5 + 5 * r0, P->S[0], P->X[0]);
        else launch_pack(*P, I, P->stage3 + 3 * r0, P->S[0]);
    }
    InstDev D;
    std::memset(&D, 0, sizeof D);
    D.K = K;
    D.time = t0;
    D.tfinal = tfinal;
    D.bc_period = 1.0;
    D.max_steps = max_steps;
    D.state = INST_RUNNING;
    D.bc_kind = SWE_BC_NONE;
    static_assert(sizeof(InstDev) <= sizeof(PokeWords), "an instance record is written in one edit");
    const unsigned long long z64 = 0ull;
    const int z32 = 0;
    if ((rc = launch_poke(*P, P->tab + k, &D, sizeof D))) return rc;
    if ((rc = launch_poke(*P, P->vmax + k, &z64, sizeof z64))) return rc;
    if ((rc = launch_poke(*P, P->flags + k, &z32, sizeof z32))) return rc;
    P->listed[(size_t)k] = 1;
    return pool_update_active(P);
}

int pool_update_active(Pool* P) {
    // wait until the staging buffer's previous upload has been consumed
    const int b = P->act_buf;
    P->act_buf ^= 1;
    SWE_HIPCHK(hipEventSynchronize(P->act_ev[b]));
    std::vector<std::pair<int64_t, int64_t>> runs;
    P->any_src = false;
    for (int k = 0; k < P->inst_cap; ++k)
        if (P->inst[(size_t)k].used && P->listed[(size_t)k]) {
            runs.emplace_back(P->inst[(size_t)k].slice0, P->inst[(size_t)k].nslices);
            Consts K;
            if (make_consts(P->inst[(size_t)k].prm, K) == SWE_OK && K.src != SRC_NONE) P->any_src = true;
        }
    std::sort(runs.begin(), runs.end());
    int n = 0;
    for (const auto& r : runs)
        for (int64_t s = r.first; s < r.first + r.second; ++s) P->act_host[b][n++] = (int)s;
    if (n) SWE_HIPCHK(hipMemcpyAsync(P->act, P->act_host[b], sizeof(int) * (size_t)n, hipMemcpyHostToDevice, P->stream));
    SWE_HIPCHK(hipEventRecord(P->act_ev[b], P->stream));
    P->nact = n;
    return SWE_OK;
}

int pool_remove(Pool* P, int k) {
    if (k < 0 || k >= P->inst_cap || !P->inst[(size_t)k].used) return fail(SWE_EINVAL, "no instance %d", k);
    InstHost& I = P->inst[(size_t)k];
    launch_fill_slices(*P, I.slice0, I.nslices, -1);
    launch_fill_bc(*P, I.bc0, I.nbc);
    const int st = INST_FREE, go = 0;
    int rc = launch_poke(*P, &P->tab[k].state, &st, sizeof st);
    if (!rc) rc = launch_poke(*P, &P->tab[k].go, &go, sizeof go);
    if (rc) return rc;
    SWE_HIPCHK(hipEventRecord(P->released, P->stream));
    P->a_slices.release(I.slice0, I.nslices);
    P->a_slots.release(I.slot0, I.nslots);
    P->a_elist.release(I.e0, I.ne);
    P->a_bc.release(I.bc0, I.nbc);
    I = InstHost{};
    P->listed[(size_t)k] = 0;
    if ((rc = pool_update_active(P))) return rc;
    bool any = false;
    for (const auto& q : P->inst) any |= q.used;
    if (!any) {
        P->mode_set = false;
        P->wmax = 0;
    }
    return SWE_OK;
}

int pool_set_field(Pool* P, int k, size_t offset, const void* src, size_t bytes) {
    return launch_poke(*P, (char*)(P->tab + k) + offset, src, bytes);
}

int pool_set_params(Pool* P, int k, const swe_params& prm) {
    Mode md;
    int rc = mode_of(prm, md);
    if (rc) return rc;
    int others = 0;
    for (int q = 0; q < P->inst_cap; ++q) others += (q != k && P->inst[(size_t)q].used) ? 1 : 0;
    if (others && !same_mode(md, P->mode)) return fail(SWE_EBADOPT, "parameters change the pool's scheme");
    Consts K;
    if ((rc = make_consts(prm, K))) return rc;
    P->mode = md;
    P->mode_set = true;
    P->inst[(size_t)k].prm = prm;
    rc = pool_set_field(P, k, offsetof(InstDev, K), &K, sizeof K);
    if (rc) return rc;
    if (K.src != SRC_NONE) P->any_src = true;    // (a list update recomputes it exactly)
    return SWE_OK;
}

int pool_set_bc(Pool* P, int k, int kind, double amp, double period, double level) {
    if (kind < SWE_BC_NONE || kind > SWE_BC_STAGES) return fail(SWE_EINVAL, "unknown bc kind %d", kind);
    if (kind == SWE_BC_INFLOW && !(period != 0.0)) return fail(SWE_EINVAL, "inflow period must be non-zero");
    const double v[3] = {amp, period, level};
    int rc = pool_set_field(P, k, offsetof(InstDev, bc_amp), v, sizeof v);
    if (rc) return rc;
    return pool_set_field(P, k, offsetof(InstDev, bc_kind), &kind, sizeof kind);
}

int pool_upload_state(Pool* P, int k, const double* un, Node* dst) {
    const InstHost& I = P->inst[(size_t)k];
    if (I.nc == 5) {
        double* st5 = P->stage5 + 5 * I.slice0 * 64;
        SWE_HIPCHK(hipMemcpyAsync(st5, un, sizeof(double) * 5 * (size_t)I.np, hipMemcpyHostToDevice, P->stream));
        launch_pack5(*P, I, st5, dst, ext_of(*P, dst));
        return SWE_OK;
    }
    double* st = P->stage3 + 3 * I.slice0 * 64;
    SWE_HIPCHK(hipMemcpyAsync(st, un, sizeof(double) * 3 * (size_t)I.np, hipMemcpyHostToDevice, P->stream));
    launch_pack(*P, I, st, dst);
    return SWE_OK;
}

int pool_download_state(Pool* P, int k, const Node* src, double* un) {
    const InstHost& I = P->inst[(size_t)k];
    if (I.nc == 5) {
        double* st5 = P->stage5 + 5 * I.slice0 * 64;
        launch_unpack5(*P, I, src, ext_of(*P, src), st5);
        SWE_HIPCHK(hipMemcpyAsync(un, st5, sizeof(double) * 5 * (size_t)I.np, hipMemcpyDeviceToHost, P->stream));
        SWE_HIPCHK(hipStreamSynchronize(P->stream));
        return SWE_OK;
    }
    double* st = P->stage3 + 3 * I.slice0 * 64;
    launch_unpack(*P, I, src, st);
    SWE_HIPCHK(hipMemcpyAsync(un, st, sizeof(double) * 3 * (size_t)I.np, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    return SWE_OK;
}

// One SSP-RK3 step of every live instance (SOURCES/main.f90:71-91): A = U^n -> B -> C -> A.
int record_dt(Pool* P) {
    const bool lam = P->mode.viscosity_type != SWE_GALERKIN;
    launch_compute_dt(*P, P->S[0], lam ? P->d0 : nullptr, 0);
    launch_finalize_dt(*P, 1);
    return SWE_OK;
}

int record_stages(Pool* P) {
    const bool lam = P->mode.viscosity_type != SWE_GALERKIN;
    Node* A = P->S[0];
    Node* B = P->S[1];
    Node* C = P->S[2];
    int rc;
    if ((rc = launch_euler_stage(*P, A, A, B, lam ? P->d0 : nullptr, 0))) return rc;
    launch_bdy(*P, B, 0);
    if ((rc = launch_euler_stage(*P, B, A, C, nullptr, 1))) return rc;
    launch_bdy(*P, C, 1);
    if ((rc = launch_euler_stage(*P, C, A, A, nullptr, 2))) return rc;
    launch_bdy(*P, A, 2);
    return SWE_OK;
}

namespace {

int record_step(Pool* P) {
    const int rc = record_dt(P);
    return rc ? rc : record_stages(P);
}

int64_t tuning_hash(const Tuning& t) {
    int64_t h = 0;
    for (int k = 0; k < KID_COUNT; ++k) h = h * 1315423911ll + (t.k[k].block * 2 + t.k[k].remap) * 4 + t.k[k].kind;
    return h;
}

// The captured step stays valid while the launch extents, the scheme and the schedules stay.
int ensure_graph(Pool* P) {
    const int64_t key[6] = {(int64_t)P->nact, (int64_t)P->ninst(), P->nbc(), (int64_t)P->wmax,
                            (int64_t)(P->mode.viscosity_type * 128 + P->mode.viscous_type * 16 + P->mode.alpha * 4 +
                                      P->mode.roundoff * 2 + (P->any_src ? 1 : 0)),
                            tuning_hash(P->tune)};
    if (P->gexec && std::equal(key, key + 6, P->graph_key)) return SWE_OK;
    if (P->gexec) {
        (void)hipGraphExecDestroy(P->gexec);
        P->gexec = nullptr;
    }
    hipGraph_t g = nullptr;
    SWE_HIPCHK(hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal));
    const int rc = record_step(P);
    const hipError_t ee = hipStreamEndCapture(P->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ee != hipSuccess) return fail(SWE_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
    const hipError_t ei = hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        P->gexec = nullptr;
        return fail(SWE_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    }
    std::copy(key, key + 6, P->graph_key);
    return SWE_OK;
}

}  // namespace

int pool_queue_steps(Pool* P, int nsteps) {
    if (nsteps <= 0 || P->nact == 0) return SWE_OK;
    int rc;
    if (P->step_form == 1 && !P->timing && persistent_applies(*P)) {
        if ((rc = launch_step_persistent(*P, nsteps))) return rc;
    } else if (P->use_graph && !P->timing) {
        if ((rc = ensure_graph(P))) return rc;
        for (int it = 0; it < nsteps; ++it) SWE_HIPCHK(hipGraphLaunch(P->gexec, P->stream));
    } else {
        for (int it = 0; it < nsteps; ++it)
            if ((rc = record_step(P))) return rc;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SWE_EHIP, "kernel launch failed: %s", hipGetErrorString(e));
    return SWE_OK;
}

int pool_snapshot(Pool* P, long long* ticket) {
    Snapshot& sn = P->snaps[P->next_ticket % 4];
    if (sn.ticket >= 0) SWE_HIPCHK(hipEventSynchronize(sn.ev));   // its buffers are about to be reused
    const int n = std::max(P->ninst(), 1);
    SWE_HIPCHK(hipMemcpyAsync(sn.tab, P->tab, sizeof(InstDev) * (size_t)n, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipMemcpyAsync(sn.flags, P->flags, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipEventRecord(sn.ev, P->stream));
    sn.n = n;
    sn.ticket = P->next_ticket++;
    *ticket = sn.ticket;
    return SWE_OK;
}

int pool_finished(Pool* P, long long ticket, int* ids, int cap, int* n) {
    *n = 0;
    if (ticket < 0 || P->snaps[ticket % 4].ticket != ticket)
        return fail(SWE_EINVAL, "snapshot %lld is not available (only the last 4 are kept)", ticket);
    Snapshot& sn = P->snaps[ticket % 4];
    SWE_HIPCHK(hipEventSynchronize(sn.ev));
    const int ni = std::min(P->ninst(), sn.n);
    bool unlisted = false;
    for (int k = 0; k < ni; ++k) {
        InstHost& I = P->inst[(size_t)k];
        // an instance admitted after the snapshot was queued is not in it (the slot may show the
        // record of an instance removed since)
        if (!I.used || I.final_known || ticket < I.since) continue;
        const InstDev& D = sn.tab[k];
        const bool over = D.state != INST_RUNNING || sn.flags[k] != 0 || !(D.time < D.tfinal && D.steps < D.max_steps);
        if (!over) continue;
        I.final_known = true;
        I.final_rec = D;
        I.final_flags = sn.flags[k];
        if (P->listed[(size_t)k]) {
            P->listed[(size_t)k] = 0;
            unlisted = true;
        }
        if (*n < cap) ids[(*n)++] = k;
    }
    return unlisted ? pool_update_active(P) : SWE_OK;
}

int pool_result(Pool* P, int k, double* un) {
    const InstHost& I = P->inst[(size_t)k];
    if (!I.final_known) return fail(SWE_EINVAL, "instance %d has not been reported finished", k);
    if (I.nc == 5) {
        double* st5 = P->stage5 + 5 * I.slice0 * 64;
        launch_unpack5(*P, I, P->S[0], P->X[0], st5, P->copy);
        SWE_HIPCHK(hipMemcpyAsync(un, st5, sizeof(double) * 5 * (size_t)I.np, hipMemcpyDeviceToHost, P->copy));
        SWE_HIPCHK(hipStreamSynchronize(P->copy));
        return SWE_OK;
    }
    double* st = P->stage3 + 3 * I.slice0 * 64;
    launch_unpack(*P, I, P->S[0], st, P->copy);
    SWE_HIPCHK(hipMemcpyAsync(un, st, sizeof(double) * 3 * (size_t)I.np, hipMemcpyDeviceToHost, P->copy));
    SWE_HIPCHK(hipStreamSynchronize(P->copy));
    return SWE_OK;
}

int pool_read_table(Pool* P) {
    const int n = std::max(P->ninst(), 1);
    SWE_HIPCHK(hipMemcpyAsync(P->tab_host, P->tab, sizeof(InstDev) * (size_t)n, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipMemcpyAsync(P->flags_host, P->flags, sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, P->stream));
    SWE_HIPCHK(hipStreamSynchronize(P->stream));
    return SWE_OK;
}

// ---- autotuner ------------------------------------------------------------------------------
// Coordinate search over the launch schedule of each kernel the scheme uses: workgroup size
// (128, 256, 512) and XCD-aware workgroup order (on/off).  Each candidate runs `steps` steps
// from a saved snapshot of the states and the instance table, timed with HIP events around that
// kernel's launches; the state and the table are restored afterwards, so tuning changes no
// result (every schedule computes the same rows with the same arithmetic).
int pool_tune(Pool* P, int steps, int trials) {
    if (P->nact == 0) return SWE_OK;
    steps = std::max(1, steps);
    trials = std::max(1, trials);
    const int n = std::max(P->ninst(), 1);
    const size_t R = (size_t)P->rows_cap;
    InstDev* tab_save = nullptr;
    int* flags_save = nullptr;
    int rc;
    if ((rc = dalloc(&tab_save, (size_t)n))) return rc;
    if ((rc = dalloc(&flags_save, (size_t)n))) {
        (void)hipFree(tab_save);
        return rc;
    }
    auto snapshot = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(P->scratch, P->S[0], sizeof(Node) * R, hipMemcpyDeviceToDevice, P->stream);
        if (e == hipSuccess && P->xscratch)
            e = hipMemcpyAsync(P->xscratch, P->X[0], sizeof(NodeX) * R, hipMemcpyDeviceToDevice, P->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(tab_save, P->tab, sizeof(InstDev) * n, hipMemcpyDeviceToDevice, P->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(flags_save, P->flags, sizeof(int) * n, hipMemcpyDeviceToDevice, P->stream);
        return e;
    };
    auto restore = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(P->S[0], P->scratch, sizeof(Node) * R, hipMemcpyDeviceToDevice, P->stream);
        if (e == hipSuccess && P->xscratch)
            e = hipMemcpyAsync(P->X[0], P->xscratch, sizeof(NodeX) * R, hipMemcpyDeviceToDevice, P->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(P->tab, tab_save, sizeof(InstDev) * n, hipMemcpyDeviceToDevice, P->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(P->flags, flags_save, sizeof(int) * n, hipMemcpyDeviceToDevice, P->stream);
        if (e == hipSuccess) e = hipMemsetAsync(P->vmax, 0, sizeof(unsigned long long) * n, P->stream);
        return e;
    };
    std::vector<int> kids = {KID_DT};
    const Mode& md = P->mode;
    const bool alpha = md.alpha && md.viscosity_type != SWE_GALERKIN;
    if (alpha) kids.push_back(KID_ALPHA);
    if (md.viscosity_type == SWE_FCT) {
        kids.push_back(KID_FCT_LO);
        kids.push_back(KID_RMINUS);
        kids.push_back(KID_APPLY);
    } else if (md.viscosity_type == SWE_ENTROPY_VISC) {
        kids.push_back(KID_FCT_LO);
        kids.push_back(KID_EV);
    } else {
        kids.push_back(KID_STAGE_LO);
    }
    const bool timing0 = P->timing;
    const std::string name0 = P->timed_name;
    hipError_t e = snapshot();
    rc = (e == hipSuccess) ? SWE_OK : fail(SWE_EHIP, "tuning snapshot: %s", hipGetErrorString(e));
    // true from the start of a trial run until its restore is queued: every exit after a good
    // snapshot gives the instances back their states, times, step counts and flags
    bool dirty = false;
    // SWE_TUNE_FAIL_AT=n makes the n-th trial fail after its steps are queued (a host-side failure
    // injection for the test of the restore path; unset or 0: never)
    const char* fa = std::getenv("SWE_TUNE_FAIL_AT");
    const int fail_at = fa ? std::atoi(fa) : 0;
    int trial_no = 0;
    const int blocks[3] = {128, 256, 512};
    for (int kid : kids) {
        if (rc) break;
        LaunchCfg best = P->tune.k[kid];
        double best_ms = std::numeric_limits<double>::infinity();
        for (int cand = 0; cand < 3 * 2 * kinds_of(kid) && !rc; ++cand) {
            const int b = blocks[cand % 3], remap = (cand / 3) % 2, kind = cand / 6;
            {
                if (!block_supported(kid, b)) continue;
                P->tune.k[kid] = LaunchCfg{b, remap, kind};
                double ms_min = std::numeric_limits<double>::infinity();
                for (int t = 0; t < trials && !rc; ++t) {
                    P->timing = true;
                    P->timed_name = kKernelNames[kid];
                    P->tk.used = 0;
                    dirty = true;
                    if ((rc = pool_queue_steps(P, steps))) break;
                    if (++trial_no == fail_at) {
                        rc = fail(SWE_EHIP, "tuning run: failure injected at trial %d (SWE_TUNE_FAIL_AT)", trial_no);
                        break;
                    }
                    if ((e = hipStreamSynchronize(P->stream)) != hipSuccess) {
                        rc = fail(SWE_EHIP, "tuning run: %s", hipGetErrorString(e));
                        break;
                    }
                    double sum = 0.0;
                    for (int q = 0; q < P->tk.used; ++q) {
                        float ms = 0.0f;
                        (void)hipEventElapsedTime(&ms, P->tk.ev[(size_t)2 * q], P->tk.ev[(size_t)2 * q + 1]);
                        sum += ms;
                    }
                    ms_min = std::min(ms_min, sum);
                    if ((e = restore()) != hipSuccess) rc = fail(SWE_EHIP, "tuning restore: %s", hipGetErrorString(e));
                    else dirty = false;
                }
                if (ms_min < best_ms) {
                    best_ms = ms_min;
                    best = P->tune.k[kid];
                }
            }
        }
        P->tune.k[kid] = best;
    }
    P->timing = timing0;
    P->timed_name = name0;
    P->tk.used = 0;
    // the step form, on whole steps: the launch sequence (captured graph when enabled) or the
    // persistent step; the first run of each form is untimed (graph capture, code load)
    if (!rc && persistent_applies(*P)) {
        hipEvent_t ev[2] = {nullptr, nullptr};
        e = hipEventCreate(&ev[0]);
        if (e == hipSuccess) e = hipEventCreate(&ev[1]);
        if (e != hipSuccess) rc = fail(SWE_EHIP, "tuning events: %s", hipGetErrorString(e));
        const int form0 = P->step_form;
        int best_form = form0;
        double best_ms = std::numeric_limits<double>::infinity();
        for (int form = 0; form <= 1 && !rc; ++form) {
            P->step_form = form;
            double ms_min = std::numeric_limits<double>::infinity();
            for (int t = 0; t <= trials && !rc; ++t) {
                dirty = true;
                if ((e = hipEventRecord(ev[0], P->stream)) != hipSuccess) {
                    rc = fail(SWE_EHIP, "tuning run: %s", hipGetErrorString(e));
                    break;
                }
                if ((rc = pool_queue_steps(P, steps))) break;
                if (++trial_no == fail_at) {
                    rc = fail(SWE_EHIP, "tuning run: failure injected at trial %d (SWE_TUNE_FAIL_AT)", trial_no);
                    break;
                }
                if ((e = hipEventRecord(ev[1], P->stream)) == hipSuccess) e = hipStreamSynchronize(P->stream);
                if (e != hipSuccess) {
                    rc = fail(SWE_EHIP, "tuning run: %s", hipGetErrorString(e));
                    break;
                }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
                if (t > 0) ms_min = std::min(ms_min, (double)ms);
                if ((e = restore()) != hipSuccess) rc = fail(SWE_EHIP, "tuning restore: %s", hipGetErrorString(e));
                else dirty = false;
            }
            if (ms_min < best_ms) {
                best_ms = ms_min;
                best_form = form;
            }
        }
        P->step_form = rc ? form0 : best_form;
        for (hipEvent_t x : ev)
            if (x) (void)hipEventDestroy(x);
    }
    if (dirty) {
        // a trial failed part-way: wait for what it queued, then restore the snapshot; a failed
        // restore is reported on its own (the tuning error is kept in the message)
        (void)hipStreamSynchronize(P->stream);
        const std::string why = last_error();
        e = restore();
        if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
        if (e != hipSuccess)
            rc = fail(SWE_EHIP, "tuning restore after a failed trial (%s): %s", why.c_str(), hipGetErrorString(e));
    }
    (void)hipStreamSynchronize(P->stream);
    (void)hipFree(tab_save);
    (void)hipFree(flags_save);
    return rc;
}

}  // namespace swe
