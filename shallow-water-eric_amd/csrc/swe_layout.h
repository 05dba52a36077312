// swe_layout.h -- data layout shared by the host engine and the kernels.
//
// A pool is one HBM arena holding any number of meshes ("instances") side by side.  Rows (nodes)
// are grouped in slices of 64, one wave each; a slice belongs to exactly one instance, so every
// wave reads its instance's constants, dt and state once, from the instance table, with scalar
// loads.  Entries of a slice are stored sliced-ELL: row r of slice s owns the slots
// sl[s].sbase + (r % 64) + 64*q for q < len[r], in the row's original CSR order.  Column indices
// are global row numbers, so a kernel never needs an instance offset.  A launch covers the list
// of slices of the instances that are stepping (rebuilt when instances join or leave), so free
// rows and finished instances cost no waves.  The single-mesh handle is a pool with one instance.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swe {

// One node of a state buffer: (h, qx, qy) and the bathymetry z, 32 bytes, so a neighbour gather
// is one aligned 32-byte load.
struct alignas(32) Node {
    double h, qx, qy, z;
};
// The two further components of the SGN model (test 13): h*eta and h*w, in arrays beside the
// Node buffers of pools that run that model.
struct alignas(16) NodeX {
    double he, hw;
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
    double dt_fixed;    // > 0: the time loop's dt instead of the CFL one (test 9, SOURCES/main.f90:68)
    double lambda_sgn;  // inputs%lambdaSGN of the SGN model (test 13)
    int type_test;
    int check_h;        // check_hmin (SOURCES/update.f90:810-822) after every stage
    int src;            // source term of type_test (SrcKind)
    int pad_;
};

enum SrcKind : int { SRC_NONE = 0, SRC_MANNING = 1, SRC_LINEAR = 2, SRC_CORIOLIS = 3, SRC_SGN = 4 };
// FLAG_STEP: a grid barrier of the persistent step waited past its bound (the step did not finish)
enum : int { FLAG_NEG_H = 1, FLAG_LAMBDA = 2, FLAG_STEP = 4 };

// One 64-row slice (one wave).
struct alignas(32) SliceInfo {
    int64_t sbase;      // first entry slot of the slice
    int64_t ebase;      // first element of the slice's compact off-diagonal list
    int inst;           // owning instance, -1 when the slice is free or not yet admitted
    int ecount;         // length of the compact off-diagonal list
    int pad_[2];
};

enum InstState : int { INST_FREE = 0, INST_RUNNING = 1, INST_DONE = 2, INST_FAILED = 3 };

// Per-instance record in device memory.  The host writes it at admission and edits single
// fields between steps; only the per-step bookkeeping kernel writes it on the device (dt, time,
// steps, go, state, bc_val), so the row kernels read it as constant data.  The dt-pass maximum
// and the error flags of an instance live in separate arrays, written with atomics.
struct alignas(64) InstDev {
    Consts K;
    double dt;              // inputs%dt of the current step
    double time;            // inputs%time
    double tfinal;          // the instance steps while time < tfinal (SOURCES/main.f90:61)
    double bc_val[3];       // amp*sin(2*pi*t/period) at the boundary times of the three stages
    double bc_amp, bc_period, bc_level;
    double* dt_log;         // optional record of dt per step (ring of dt_cap entries)
    long long steps;        // steps taken
    long long max_steps;    // and the limit on them
    int dt_cap;
    int state;              // InstState
    int go;                 // 1 while the instance takes the current step
    int bc_kind;            // swe_bc_kind
};

// A Dirichlet overwrite of bdy: node (global row), owning instance, component (0 h, 1 qx, 2 qy)
// and the index of its value for fixed-value boundaries.
struct BcEntry {
    int node, inst, comp, vidx;
};

// Column offsets that do not fit 16 bits are marked with kColFar and read from the 32-bit columns.
constexpr int kColFar = -32768;

// Read-only view of the arena's mesh arrays handed to every row kernel.
struct Mesh {
    const SliceInfo* __restrict__ sl;
    const uint8_t* __restrict__ len;    // entries of each row (0 for padding rows)
    const int* __restrict__ col;        // global column of each slot
    const int16_t* __restrict__ dcol;   // column - row of each slot, kColFar where that does not fit
    const double2* __restrict__ c;      // (cx, cy) of each slot
    const double* __restrict__ m;       // lumped mass of each row
    const uint16_t* __restrict__ elist; // per slice: slot offsets (q*64 + lane) of off-diagonal entries
    const uint8_t* __restrict__ dpos;   // position q of the diagonal in each row
    const int* __restrict__ act;        // the slices a launch covers, ascending (one wave each)
    int nact;                           // number of them
    int wmax;                           // longest row of the resident instances
};

}  // namespace swe
