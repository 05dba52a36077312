// swe_engine.h -- host side of the MI355X step: the pool (an HBM arena of instances), its
// allocators, the launch wrappers of the kernels and the step sequence.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "swe_amd.h"
#include "swe_layout.h"

namespace swe {

// ---- errors --------------------------------------------------------------------------------
int fail(int code, const char* fmt, ...);
const char* last_error();
void clear_error();

#define SWE_HIPCHK(x)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return ::swe::fail(SWE_EHIP, "%s failed: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// Longest row: the dt pass keeps (2 + w)*64 doubles per wave in LDS.
constexpr int kMaxRow = 30;

// ---- launch schedules (chosen per kernel by the autotuner) -----------------------------------
enum KernelId : int { KID_DT = 0, KID_ALPHA, KID_STAGE_LO, KID_FCT_LO, KID_RMINUS, KID_APPLY, KID_EV, KID_COUNT };
extern const char* const kKernelNames[KID_COUNT];

struct LaunchCfg {
    int block = 256;    // threads per workgroup (whole waves; 4 slices at 256)
    int remap = 1;      // XCD-aware workgroup order
    int kind = 0;       // schedule kind (k_fct_lo in stages 2-3: 0 lambda per row inline,
                        // 1 lambda entry-parallel in LDS); 0 for the other kernels
};
// Schedule kinds a kernel has (1 or 2).
int kinds_of(int kid);
struct Tuning {
    LaunchCfg k[KID_COUNT];   // defaults: 256 threads, XCD-aware order (the tuner refines per mesh)
};
// Workgroup sizes a kernel can be launched with.
bool block_supported(int kid, int block);

// ---- first-fit allocator over [0, cap) --------------------------------------------------------
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;   // start -> length
    int64_t cap = 0;
    void init(int64_t c);
    int64_t alloc(int64_t n);           // start, or -1
    void release(int64_t start, int64_t n);
    int64_t high_water() const;         // end of the last allocated range
    int64_t largest_free() const;
};

// ---- host layout of one mesh ------------------------------------------------------------------
struct HostLayout {
    int np = 0, nnz = 0, np_pad = 0, nsl = 0, wmax = 0;
    int64_t nslot = 0;
    std::vector<int> perm, iperm, len, dpos, col;     // col: local internal numbering
    std::vector<int16_t> dcol;                        // col - row, or kColFar
    std::vector<int64_t> sptr, eoff, csr2slot;        // local slots / list offsets
    std::vector<double2> c;
    std::vector<double> m, z;
    std::vector<uint16_t> elist;
    std::vector<int> lists[3];                        // Dirichlet lists, internal numbering
};
int validate_mesh(const swe_mesh_desc& d);
int build_layout(const swe_mesh_desc& d, int sort_window, HostLayout& L);
// Rows sorted by length inside windows of this many rows: SWE_SORT_WINDOW, else 64.
int sort_window_default();
// SWE_REORDER: 0 keeps the caller's numbering as the base order, 1 (default) renumbers for locality
// when the caller's numbering scatters neighbours (mean |i-j| above 2 sqrt(np)), 2 always.
int reorder_mode();
int make_consts(const swe_params& p, Consts& K);

// ---- the pool ---------------------------------------------------------------------------------
struct InstHost {
    bool used = false;
    bool final_known = false;         // reported finished by a snapshot: its rows no longer change
    InstDev final_rec{};              // that snapshot's record
    int final_flags = 0;
    long long since = 0;              // first table snapshot ticket that shows this instance
    int np = 0, nnz = 0, np_pad = 0, wmax = 0;
    int64_t slice0 = 0, nslices = 0, slot0 = 0, nslots = 0, e0 = 0, ne = 0, bc0 = 0, nbc = 0;
    int nh = 0, nux = 0, nuy = 0;
    int nc = 3;                       // components of the state: 3, or 5 for the SGN model
    swe_params prm{};
};

struct Mode {
    int viscosity_type = SWE_VISCOUS, viscous_type = 2, alpha = 0, roundoff = 1;
    int sgn = 0;                      // the 5-component SGN model (type_test 13)
};
int mode_of(const swe_params& p, Mode& m);
bool same_mode(const Mode& a, const Mode& b);

// Table snapshots taken after queued steps (pinned copies, completion event, ticket).
struct Snapshot {
    InstDev* tab = nullptr;
    int* flags = nullptr;
    hipEvent_t ev = nullptr;
    long long ticket = -1;
    int n = 0;                        // records copied
};

struct TimedKernel {
    std::vector<hipEvent_t> ev;   // start/stop pairs
    int used = 0;
};

struct Pool {
    int device = 0;
    hipStream_t stream = nullptr;     // steps and every ordered edit of the tables
    hipStream_t copy = nullptr;       // bulk uploads of admitted instances
    hipEvent_t released = nullptr;    // recorded on `stream` when ranges are given back
    bool mode_set = false;
    Mode mode{};
    int sort_window = 64;
    // capacities and allocators
    int64_t rows_cap = 0, slots_cap = 0, elist_cap = 0, bc_cap = 0;
    int inst_cap = 0;
    RangeAlloc a_slices, a_slots, a_elist, a_bc;
    std::vector<InstHost> inst;
    int wmax = 0;
    // device arrays: mesh
    SliceInfo* sl = nullptr;
    uint8_t* len = nullptr;           // row lengths and diagonal positions (rows hold at most kMaxRow)
    uint8_t* dpos = nullptr;
    int* perm = nullptr;              // internal row -> caller's node (local numbering)
    double* m = nullptr;
    double* z = nullptr;
    int* col = nullptr;
    int16_t* dcol = nullptr;          // column offsets (read instead of col where they fit)
    double2* c = nullptr;
    uint16_t* elist = nullptr;
    // the slices the step kernels cover: those of the instances that are (or may be) stepping
    int* act = nullptr;
    int nact = 0;
    int* act_host[2] = {nullptr, nullptr};   // pinned staging, used alternately
    hipEvent_t act_ev[2] = {nullptr, nullptr};
    int act_buf = 0;
    std::vector<char> listed;                // per instance: its slices are in the list
    bool any_src = false;                    // a listed instance has a source term (type_test)
    // states and work arrays
    Node* S[3] = {nullptr, nullptr, nullptr};   // S[0] holds U^n between steps
    Node* scratch = nullptr;
    // SGN pools only (allocated at the first SGN admission): h*eta, h*w beside S[0], S[1], S[2] and
    // scratch, and a 5-doubles-per-row staging buffer for un(5,np)
    NodeX* X[4] = {nullptr, nullptr, nullptr, nullptr};
    double* stage5 = nullptr;
    NodeX* xscratch = nullptr;        // the tuner's copy of X[0]
    double* ulo_h = nullptr;          // FCT low-order solution: depth
    double2* ulo_q = nullptr;         // and discharge
    double *d0 = nullptr, *ework = nullptr;
    double2* dmu = nullptr;           // first-order viscosities (dL, muL) of the FCT passes, per slot
    double* dL() const { return (double*)dmu; }   // the same memory as a per-slot scratch array
    double *alpha = nullptr, *res = nullptr, *Rm = nullptr, *nwork = nullptr, *stage3 = nullptr;
    // instances and boundaries
    InstDev* tab = nullptr;
    unsigned long long* vmax = nullptr;
    int* flags = nullptr;
    BcEntry* bc = nullptr;
    double* bcv = nullptr;            // Dirichlet values: stage s of entry v at bcv[s * bcv_stride + v]
    int64_t bcv_stride = 1;
    InstDev* tab_host = nullptr;      // pinned read-back of the table
    int* flags_host = nullptr;        // pinned read-back of the flags
    Snapshot snaps[4];
    long long next_ticket = 0;
    // schedules, graph, timing
    Tuning tune;
    int step_form = 0;                // swe_step_form: 0 launch sequence (graph), 1 persistent (fct)
    bool coop_ok = false;             // the device takes cooperative launches
    int num_cu = 1;
    unsigned* bar = nullptr;          // the persistent step's barrier word and abort word
    bool use_graph = false;
    hipGraphExec_t gexec = nullptr;
    int64_t graph_key[6] = {-1, -1, -1, -1, -1, -1};
    bool timing = false;
    std::string timed_name;
    TimedKernel tk;

    int nrows() const { return (int)(a_slices.high_water() * 64); }
    int ninst() const;                // 1 + highest used instance index
    int64_t nbc() const { return a_bc.high_water(); }
};

struct PoolCaps {
    int64_t rows = 0, slots = 0, elist = 0, bc = 0;
    int instances = 1;
};
size_t pool_bytes(const PoolCaps& caps);
int pool_create(const PoolCaps& caps, Pool** out);
void pool_destroy(Pool* P);
// Admit a mesh: bulk data goes up on the copy stream; the instance starts stepping at the next
// step queued after this call.  un may be null (state set later).  keep receives the layout
// when the caller needs its maps (csr2slot, lists).
int pool_add(Pool* P, const swe_mesh_desc& d, const swe_params& prm, const double* un, double t0,
             double tfinal, long long max_steps, int* id, HostLayout* keep);
// Make instance k live once its mesh arrays are in place (layout: its ranges and sizes); with
// packed_state, its state is in stage3 in the caller's numbering.  The instance joins the next step.
int pool_activate(Pool* P, int k, const InstHost& layout, const swe_params& prm, bool packed_state, double t0,
                  double tfinal, long long max_steps);
int pool_remove(Pool* P, int id);
int pool_set_params(Pool* P, int id, const swe_params& prm);
int pool_set_bc(Pool* P, int id, int kind, double amp, double period, double level);
int pool_set_field(Pool* P, int id, size_t offset, const void* src, size_t bytes);
int pool_upload_state(Pool* P, int id, const double* un, Node* dst);
// SGN pools: make the h*eta/h*w arrays at the first SGN admission; SGN instances take h lists only.
int ensure_sgn_buffers(Pool* P);
int check_sgn_lists(const Mode& md, int nux, int nuy);
int pool_download_state(Pool* P, int id, const Node* src, double* un);
int pool_queue_steps(Pool* P, int nsteps);
// The parts of one step, launched on the compute stream: the dt pass and its bookkeeping, then the
// three stages with bdy.
int record_dt(Pool* P);
int record_stages(Pool* P);
// Rebuild the launch list from the instances marked listed (ordered on the compute stream).
int pool_update_active(Pool* P);
int pool_read_table(Pool* P);         // sync, then tab_host / flags_host hold the table
// Queue a copy of the table after the work queued so far; returns its ticket.
int pool_snapshot(Pool* P, long long* ticket);
// Wait for a snapshot (not for later work) and list the instances it shows done or failed; they
// leave the launch list and their final records are kept for pool_result.
int pool_finished(Pool* P, long long ticket, int* ids, int cap, int* n);
// Final state of an instance reported by pool_finished, read on the copy stream while steps run.
int pool_result(Pool* P, int id, double* un);
int pool_tune(Pool* P, int steps, int trials);

// ---- layout of a mesh assembled on the device (swe_setup_device.hip) --------------------------
// The same internal numbering and sliced layout build_layout makes on the host (rows sorted by
// length inside 64-row windows of the caller's numbering), computed on the GPU from an assembly.
struct DeviceLayout {
    int np = 0, nsl = 0, wmax = 0;
    int64_t nnz = 0, nslot = 0, ne = 0;
    int* perm = nullptr;       // internal row -> caller's node
    int* iperm = nullptr;      // caller's node -> internal row
    int* len = nullptr;        // entries of each internal row (np_pad)
    int64_t* sptr = nullptr;   // slot offset of each slice (nsl + 1)
    int64_t* eoff = nullptr;   // compact-list offset of each slice (nsl + 1)
    hipStream_t stream = nullptr;   // the assembly's stream: the buffers are allocated and freed on it
};
int build_layout_device(const swe_assembly* a, DeviceLayout& DL);
void free_device_layout(DeviceLayout& DL);
// Admit an assembled mesh with its device layout; bath(np) and the 1-based Dirichlet lists are host
// arrays; csr2slot (may be null) receives the slot of each CSR entry, relative to the instance.
// compute_errors on the device (swe_errors.hip): element list and coordinates of one instance
struct ErrEval;
int errors_create(Pool* P, int k, int me, const int* jj1, const double* rr, ErrEval** out);
int errors_eval(Pool* P, const ErrEval* E, const Node* U, int type_test, double t, double g, double htiny,
                double out[12]);
int errors_flow(Pool* P, const ErrEval* E, const Node* U, double g, double out[4]);
void errors_destroy(ErrEval* E);

int pool_add_assembled(Pool* P, const swe_assembly* a, const DeviceLayout& DL, const double* bath,
                       const int* const lists[3], const int nlist[3], const swe_params& prm, const double* un,
                       double t0, double tfinal, long long max_steps, int* id, int64_t* csr2slot);

// ---- launch wrappers (swe_kernels.hip) --------------------------------------------------------
Mesh mesh_view(const Pool& P);
void launch_compute_dt(Pool& P, const Node* U, double* d0, int all);
void launch_finalize_dt(Pool& P, int advance);
int launch_euler_stage(Pool& P, const Node* U, const Node* Uo, Node* out, const double* d0, int stage);
void launch_bdy(Pool& P, Node* u, int stage);
// The persistent step form: whether it covers the pool's scheme (fct, three components) on this
// device, and nsteps steps of every listed instance in cooperative launches.
bool persistent_applies(const Pool& P);
int launch_step_persistent(Pool& P, int nsteps);
void launch_dij_full(Pool& P, const Node* U, double* dij);
void launch_pack(Pool& P, const InstHost& I, const double* un3, Node* dst);
void launch_unpack(Pool& P, const InstHost& I, const Node* src, double* un3, hipStream_t stream = nullptr);
// The same for un(5,np) of the SGN model: (h, qx, qy) into a Node buffer, (h*eta, h*w) into its X.
void launch_pack5(Pool& P, const InstHost& I, const double* un5, Node* dst, NodeX* xdst);
void launch_unpack5(Pool& P, const InstHost& I, const Node* src, const NodeX* xsrc, double* un5,
                    hipStream_t stream = nullptr);
// X array paired with a state buffer (S[0..2], scratch), or null when the pool has none.
NodeX* ext_of(const Pool& P, const Node* s);
void launch_perm_nodes(Pool& P, const InstHost& I, const double* src, double* dst);
void launch_csr_to_sell(Pool& P, const double* src, const int64_t* map, double* dst, int64_t nnz);
void launch_sell_to_csr(Pool& P, const double* src, const int64_t* map, double* dst, int64_t nnz);
void launch_fctpos(Pool& P, const double* f, const double* us, double htiny, double* lij);
void launch_fill_slices(Pool& P, int64_t s0, int64_t n, int inst);
void launch_fill_bc(Pool& P, int64_t b0, int64_t n);
// Queue a write of up to sizeof(PokeWords) bytes (whole 4-byte words) on the compute stream.
constexpr int kPokeWords = 128;
struct PokeWords {
    uint32_t w[kPokeWords];
};
int launch_poke(Pool& P, void* dst, const void* src, size_t bytes);
// Achievable HBM bandwidth: best of `reps` streaming reads and copies of `bytes` (GB/s).
int stream_bandwidth(size_t bytes, int reps, double* read_gbs, double* copy_gbs);
int launch_fctpos_csr(int np, int nnz, const int* ia, const int* ja, const double* f, const double* us,
                      double htiny, double* lij);
// The same with R- in the library's fixed device array (rows up to kNoAllocRows), queued on
// `stream` without allocating or waiting.
constexpr int kNoAllocRows = 1 << 22;
int launch_fctpos_csr_noalloc(int np, int nnz, const int* ia, const int* ja, const double* f, const double* us,
                              double htiny, double* lij, hipStream_t stream);
int launch_fctgen_csr(int np, const int* ia, const int* ja, const double* f, const double* maxn, const double* minn,
                      const double* dg, double* u);

}  // namespace swe
