/* swe_amd.h -- C-ABI of the MI355X (gfx950) explicit shallow-water step.
 *
 * The library replaces the explicit time step of t-bltg/shallow-water-eric: MODULE update
 * (SOURCES/update.f90) and MODULE fct (SOURCES/fct.f90), with the pointwise physics they call in
 * SOURCES/condlim.f90, and the per-step work of the driver's time loop (SOURCES/main.f90:61-92:
 * COMPUTE_DT, the SSP-RK3 combinations and bdy).  Every entry point takes plain pointers and
 * sizes; arrays keep the reference's layout: 1-based CSR (ia(np+1), ja(nnz), columns ascending,
 * diagonal included, as st_csr builds it), state un(3,np) in Fortran order (h, qx, qy per node),
 * c_ij as two arrays cij(1)%aa and cij(2)%aa.  With type_test 13 (the SGN model) every state
 * array is un(5,np) (h, qx, qy, h*eta, h*w), as SOURCES/main.f90:22-26 sets syst_size; such
 * solvers take h Dirichlet lists only (bdy of h*eta and h*w, main.f90:184-187, uses sol_anal(4:5),
 * which the reference defines only at t <= 1e-10) and no FCT entry points.
 *
 * All arithmetic is IEEE binary64 in the reference's evaluation order; results agree bit for bit
 * with the reference CPU step except where a source term calls pow (Manning friction), where
 * they agree to rounding.
 *
 * Threading: a handle (or batch) owns its device buffers and its HIP streams; use it from one host
 * thread at a time.  No host pointer is kept after a call returns.
 * Errors: every call returns a swe_status; swe_last_error() gives a message for the calling
 * thread.
 */
#ifndef SWE_AMD_H
#define SWE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum swe_status {
    SWE_OK = 0,
    SWE_ENEG_H = 1,     /* 'Min h<0, STOP' of check_hmin (SOURCES/update.f90:816-820) */
    SWE_ELAMBDA = 2,    /* 'BUG in compute_lambda_vacc' (SOURCES/condlim.f90:934-937) */
    SWE_EBADOPT = 3,    /* option strings the reference rejects (SOURCES/update.f90:199,217,243,273) */
    SWE_EHIP = 4,       /* a HIP runtime call failed */
    SWE_EINVAL = 5,     /* malformed arguments (sizes, indices, null pointers) */
    SWE_ENOMEM = 6,     /* device allocation failed */
    SWE_EFULL = 7       /* a batch has no room for this mesh now (retry after removing instances) */
} swe_status;

typedef enum swe_viscosity {      /* inputs%viscosity_type (SOURCES/read_data.f90:57-58) */
    SWE_GALERKIN = 0,
    SWE_VISCOUS = 1,
    SWE_ENTROPY_VISC = 2,
    SWE_FCT = 3
} swe_viscosity;

typedef enum swe_bc_kind {        /* Dirichlet values written by bdy after each stage */
    SWE_BC_NONE = 0,              /* no Dirichlet nodes */
    SWE_BC_WALLS = 1,             /* qx = 0 on the ux list, qy = 0 on the uy list */
    SWE_BC_INFLOW = 2,            /* walls, plus h = max(level - z + amp*sin(2*pi*t/period), 0) on the h list */
    SWE_BC_VALUES = 3,            /* fixed values given per listed node by swe_set_bc_values */
    SWE_BC_STAGES = 4             /* values per stage given with each step by swe_step_finish */
} swe_bc_kind;

typedef struct swe_params {
    double gravity;       /* inputs%gravity */
    double htiny;         /* inputs%htiny = epsilon*max_water_h (SOURCES/condlim.f90:233) */
    double cfl;           /* inputs%CFL */
    double max_water_h;   /* max_water_h (entropy-residual scaling, SOURCES/update.f90:866) */
    double mannings;      /* inputs%mannings (tests 9 and 10) */
    int type_test;        /* inputs%type_test, 1..13: 9/10 Manning friction, 11 linear friction, 12 Coriolis,
                             13 the SGN model (5 components, mSGN_RHS); test 9 also fixes htiny = 1e-3 in
                             every call and dt = 2.5e-2 in the time loop (SOURCES/main.f90:66-69).
                             Outside 1..13 compute_lambda_vacc STOPs (SOURCES/condlim.f90:952-954): such
                             values give SWE_EBADOPT, and so does test 13 with 'entropy_visc' or 'fct',
                             whose entropy_residual reads Entprime(4:5,:) that the reference never sets
                             (SOURCES/update.f90:853-862). */
    int viscosity_type;   /* swe_viscosity */
    int viscous_type;     /* 1 = 'type1' (smb_1), 2 = 'type2' (smb_2 family) */
    int alpha_limit;      /* inputs%if_alpha_limit */
    int roundoff_fix;     /* if_roundoff_fix (SOURCES/update.f90:15; the reference fixes .TRUE.) */
    double lambda_sgn;    /* inputs%lambdaSGN (SOURCES/read_data.f90:92-95), used by test 13 only */
} swe_params;

typedef struct swe_mesh_desc {
    int np;               /* mesh%np */
    int nnz;              /* SIZE(ja) */
    const int *ia;        /* 1-based row pointers, np+1 */
    const int *ja;        /* 1-based column indices, nnz */
    const double *cx;     /* cij(1)%aa, nnz */
    const double *cy;     /* cij(2)%aa, nnz */
    const double *lumped; /* lumped mass, np */
    const double *bath;   /* bathymetry z, np */
    int n_h_dir, n_ux_dir, n_uy_dir;     /* sizes of h_js_D, ux_js_D, uy_js_D */
    const int *h_dir, *ux_dir, *uy_dir;  /* 1-based node lists (SOURCES/mesh.f90:25-34) */
} swe_mesh_desc;

typedef struct swe_handle swe_handle;

/* Message describing the last error of the calling thread ("" if none). */
const char *swe_last_error(void);

/* P1 set-up of construct_matrices (SOURCES/update.f90:18-132) with the geometry of
 * gauss_points_2d_p1 (SFEMaNS/GAUSS_POINTS_OBJET/gauss_points_2d_p1.f90:106-132) and the pattern
 * of st_csr (SFEMaNS/FEMSUB/st_sparsekit.f90:788-863).  jj is 1-based jj(3,me); rr is rr(2,np).
 * swe_assemble_nnz returns SIZE(ja); swe_assemble fills ia, ja (1-based), diag (1-based),
 * mass%aa, lumped, cij(1)%aa, cij(2)%aa. */
long swe_assemble_nnz(int np, int me, const int *jj);
int swe_assemble(int np, int me, const int *jj, const double *rr, int *ia, int *ja, int *diag,
                 double *mass, double *lumped, double *cx, double *cy);

/* The same set-up on the GPU, bitwise equal to swe_assemble: swe_assemble_device builds the
 * pattern and the matrices in device memory and returns their number of entries (SIZE(ja));
 * swe_assembly_get copies them out in swe_assemble's layout (any pointer may be NULL);
 * swe_assembly_destroy frees them.  A node may belong to at most 32 elements. */
typedef struct swe_assembly swe_assembly;
int swe_assemble_device(int np, int me, const int *jj, const double *rr, swe_assembly **out, long long *nnz);
int swe_assembly_get(swe_assembly *a, int *ia, int *ja, int *diag, double *mass, double *lumped, double *cx,
                     double *cy);
int swe_assembly_destroy(swe_assembly *a);

/* Create a solver over one mesh: copies the mesh arrays to the device.  Replaces the module
 * state built by construct_matrices (SOURCES/update.f90:18-132) and construct_mesh.
 * The environment variable SWE_SORT_WINDOW (default 64) sets the window of rows inside which
 * the internal numbering sorts rows by length; SWE_REORDER (default 1) renumbers rows internally by
 * reverse Cuthill-McKee when the caller's numbering scatters neighbours (mean |i-j| above
 * 2*sqrt(np)), 0 never, 2 always.  These are layout choices; results do not depend on them. */
int swe_create(const swe_mesh_desc *mesh, const swe_params *prm, swe_handle **out);
/* swe_create over a mesh assembled by swe_assemble_device, without taking the matrices to the
 * host: the solver's layout is built on the GPU from the assembly (rows sorted by length inside
 * 64-row windows of a base order: the caller's numbering, or under SWE_REORDER -- see swe_create --
 * the nodes sorted along a Hilbert curve of their coordinates), bath and the Dirichlet
 * lists (1-based, as in swe_mesh_desc) are copied from the host.  The solver does not keep the
 * assembly; results are bitwise those of swe_create on the same arrays.  This replaces, together
 * with swe_assemble_device, construct_matrices (SOURCES/update.f90:18-132, called at
 * SOURCES/main.f90:21) for runs that keep their matrices on the device. */
int swe_create_from_assembly(const swe_assembly *a, const double *bath, int n_h_dir, const int *h_dir,
                             int n_ux_dir, const int *ux_dir, int n_uy_dir, const int *uy_dir,
                             const swe_params *prm, swe_handle **out);
/* compute_errors (SOURCES/main.f90:198-263) on the device state of a solver: sol_anal
 * (SOURCES/condlim.f90:311-831) at the Gauss points of r_gauss (main.f90:303-322) and the sums of
 * ns_anal_l1 / ns_anal_0 (SFEMaNS/FEMSUB_OBJET/fem_tn.f90:1679-1716, 1513-1553) for the analytic
 * tests 3 to 7, 11, 12 and 13 (the solver's type_test; others give SWE_EBADOPT; test 13 takes h, qx,
 * qy of its five components; for the steady bump flows 3, 6 and 7 the device's acos, cos and pow, and
 * for test 13 its cosh, may differ from the C library's in the last bit of a summand).
 * swe_errors_create takes
 * the element list jj(3,me) (1-based, the solver's node numbering) and rr(2,np); the handle must
 * outlive the evaluator.  swe_compute_errors writes out[12]: for k = h, qx, qy in turn the L1 error
 * of u(k,:) - exact, its L1 norm (0 - exact), the L2 error and the L2 norm.  Each sum is accumulated
 * in double-double over a fixed partition (the exactly rounded sum, the same on every run), not
 * sequentially as the reference does: the two differ by at most the sequential sum's rounding
 * error.  The reference prints err/norm of h (main.f90:214,221). */
typedef struct swe_errors swe_errors;
int swe_errors_create(swe_handle *h, int me, const int *jj, const double *rr, swe_errors **out);
int swe_compute_errors(swe_errors *e, double time, double *out);
/* The diagnostic the time loop writes to unit 10 after every step of tests 1, 2 and 10
 * (SOURCES/main.f90:94-99, replacing its three host calls of ns_l1, SFEMaNS/FEMSUB_OBJET/fem_tn.f90:
 * 2119-2153, on the device state): out[4] = {q1, q2, q3, 0.5*(q1+q2)/(gravity*q3)}, the L1 norms
 * (3-point P1 quadrature of |u_h|) of qx, qy and h*sqrt(h).  Any type_test; the sums are accumulated
 * as swe_compute_errors' are (exactly rounded, not sequential). */
int swe_ns_l1_ratio(swe_errors *e, double *out);
int swe_errors_destroy(swe_errors *e);
int swe_destroy(swe_handle *h);
int swe_set_params(swe_handle *h, const swe_params *prm);

/* Dirichlet program applied by the device-resident steps (bdy, SOURCES/main.f90:176-188). */
int swe_set_bc(swe_handle *h, int kind, double amp, double period, double level);
int swe_set_bc_values(swe_handle *h, const double *h_vals, const double *ux_vals, const double *uy_vals);

/* compute_dij (SOURCES/update.f90:291-326): dij%aa at state un(3,np) (host arrays). */
int swe_compute_dij(swe_handle *h, const double *un, double *dij);

/* COMPUTE_DT (SOURCES/main.f90:166-174): dt = CFL/MAXVAL(ABS(dij%aa(diag))/lumped).
 * un == NULL uses the device-resident state. */
int swe_compute_dt(swe_handle *h, const double *un, double *dt);

/* euler(un, unext) (SOURCES/update.f90:134-289): one forward-Euler stage with time step dt on
 * host arrays un(3,np) -> unext(3,np).  Per-stage drop-in form; copies over PCIe each call. */
int swe_euler(swe_handle *h, const double *un, double *unext, double dt);

/* FCT_positivity (SOURCES/fct.f90:5-42) on the handle's CSR pattern: unext_scaled(np) is
 * (lumped/dt)*h of the low-order solution, f(nnz) the antidiffusive fluxes, lij(nnz) out. */
int swe_fct_positivity(swe_handle *h, const double *unext_scaled, const double *f, double *lij);

/* FCT_positivity (SOURCES/fct.f90:5-42) on any 1-based CSR matrix ia(np+1), ja(nnz) without a
 * solver handle: the MODULE fct entry point of the Fortran binding. */
int swe_fct_positivity_csr(int np, int nnz, const int *ia, const int *ja, const double *unext_scaled,
                           const double *f, double htiny, double *lij);

/* swe_fct_positivity_csr without any allocation, for callers that cannot allocate device memory
 * (code under HIP stream capture, tight serving loops, harnesses that forbid hipMalloc).  It has
 * no counterpart in the reference.  Every pointer is DEVICE memory: ia(np+1), ja(nnz) 1-based,
 * unext_scaled(np), f(nnz) in, lij(nnz) out; the kernels are queued on `stream` (a hipStream_t,
 * NULL for the null stream) and the call returns without waiting.  R- of the rows lives in a
 * fixed device array of the library (4,194,304 rows, 32 MiB per GPU, held until the process
 * ends): np above that is refused with SWE_EINVAL.  The array is shared by every call on one GPU,
 * so calls that overlap on one GPU race; each call writes all of R- before reading it.  The
 * pattern is not checked on the host: entries with a column outside 1..np get lij = NaN and
 * row ranges are clipped to 1..nnz.  Results are bitwise those of swe_fct_positivity_csr. */
int swe_fct_positivity_csr_noalloc(int np, int nnz, const int *ia, const int *ja, const double *unext_scaled,
                                   const double *f, double htiny, double *lij, void *stream);

/* FCT_generic (SOURCES/SAFE/fct.f90:3-43), the two-sided limiter the reference keeps but does not
 * call: with local bounds maxn(np), minn(np), fluxes f(nnz) on a 1-based CSR matrix and the lumped
 * mass dg(np), unext += (SUM_j l_ij f_ij)/dg, l_ij = MIN(R+_i, R-_j) for f_ij >= 0 and
 * MIN(R-_i, R+_j) otherwise. */
int swe_fct_generic_csr(int np, int nnz, const int *ia, const int *ja, double *unext, const double *maxn,
                        const double *minn, const double *f, const double *dg);

/* Device-resident state. */
int swe_set_state(swe_handle *h, const double *un);
int swe_get_state(swe_handle *h, double *un);

/* Simulation time of the handle (inputs%time); it starts at 0 and each step adds its dt. */
int swe_set_time(swe_handle *h, double t);
int swe_get_time(swe_handle *h, double *t);

/* nsteps SSP-RK3 steps on the device-resident state, as in SOURCES/main.f90:71-91 with a
 * fixed step count: COMPUTE_DT, three euler stages, the combinations (3*uo+ui)/4 and
 * (uo+2*ui)/3, and bdy after each stage.  If time is not NULL the run starts at *time and
 * *time receives the end time; dts (may be NULL) receives dt of each step.  Returns
 * SWE_ENEG_H / SWE_ELAMBDA if any stage tripped those checks. */
int swe_run(swe_handle *h, int nsteps, double *time, double *dts);

/* One step in two calls, for Dirichlet values the host computes per stage (bdy with sol_anal at
 * t+dt, t+dt/2 and t+dt, SOURCES/main.f90:78,84,90), which need dt first.  swe_step_begin runs
 * COMPUTE_DT on the device-resident state and returns dt (the handle's time becomes t+dt at the
 * end of the step, as main.f90:91 sets it); swe_step_finish runs the three stages.  With the
 * SWE_BC_STAGES program, vals holds the values of the three stages one after the other, each
 * as swe_set_bc_values orders them (h list, ux list, uy list); with other programs vals may be
 * NULL.  Returns like swe_run. */
int swe_step_begin(swe_handle *h, double *dt);
int swe_step_finish(swe_handle *h, const double *vals);

/* The same steps queued without waiting for completion (dt, the time and the inflow values are
 * evaluated on the device).  swe_sync waits for the queued work and reports the checks. */
int swe_run_async(swe_handle *h, int nsteps);
int swe_sync(swe_handle *h);

/* The handle's HIP stream (hipStream_t), for callers that time or order their own work. */
void *swe_stream(swe_handle *h);

/* Per-launch timing of one kernel (by name, e.g. "k_fct_lo"; NULL stops timing) with HIP events
 * recorded around each of its launches on the handle's stream.  swe_kernel_timing returns the
 * number of launches since timing was enabled and their summed duration in milliseconds. */
int swe_enable_kernel_timing(swe_handle *h, const char *kernel);
int swe_kernel_timing(swe_handle *h, const char *kernel, int *launches, double *ms);

/* The achievable HBM bandwidth of the current device, the ceiling the kernels' roofline fractions
 * are quoted against besides the published peak: the best of `reps` streaming reads and streaming
 * copies of `bytes` bytes (16-byte non-temporal loads and stores), in GB/s of bytes moved.  It
 * allocates and frees two buffers of `bytes`.  No reference counterpart (a measurement utility). */
int swe_stream_bandwidth(unsigned long long bytes, int reps, double *read_gbs, double *copy_gbs);

/* Launch schedules: per kernel, the workgroup size (128, 256 or 512) and whether workgroups are
 * ordered XCD-aware (1) or in dispatch order (0).  Kernels, in this order: k_compute_dt, k_alpha,
 * k_stage_lo, k_fct_lo, k_fct_rminus, k_fct_apply, k_ev_update.  swe_tune times every schedule
 * of each kernel the scheme uses on the handle's own mesh and state (the state and time are
 * restored afterwards) and keeps the fastest; results do not depend on the schedule.  The
 * first-order and SGN stages (k_stage_lo) take their workgroup size and order from the same entry.
 * SWE_TUNE_FAIL_AT=n (environment, for the test of the restore path) makes the n-th trial fail
 * after its steps are queued; unset or 0 never. */
#define SWE_NUM_KERNELS 7
int swe_tune(swe_handle *h, int steps);
int swe_get_schedule(swe_handle *h, int *blocks, int *remap);
int swe_set_schedule(swe_handle *h, const int *blocks, const int *remap);
/* The schedule kind per kernel (same order), a third axis of swe_tune: k_fct_lo in stages 2 and 3
 * has kind 0 (lambda_ij evaluated per entry inside each lane's row loop) and kind 1 (lambda_ij
 * evaluated first entry-parallel, one off-diagonal entry per lane, into LDS, then read by the row
 * loop); every other kernel has kind 0 only.  swe_set_schedule_kinds returns SWE_EINVAL for a kind
 * a kernel does not have.  Results do not depend on the kind. */
int swe_get_schedule_kinds(swe_handle *h, int *kinds);
int swe_set_schedule_kinds(swe_handle *h, const int *kinds);
/* Queue steps as one captured HIP graph per step (1, the default) or as individual launches (0).
 * Timed kernels (swe_enable_kernel_timing) always launch individually. */
int swe_use_graph(swe_handle *h, int on);
/* How swe_run (and swe_batch_step*) queue steps.  SWE_STEP_LAUNCHES (the default): the passes of a
 * step as a sequence of kernel launches, replayed as one captured graph per step when graphs are
 * on.  SWE_STEP_PERSISTENT: the fct scheme's steps in cooperative launches of one kernel that runs
 * every pass, with grid-wide barriers where the launch sequence has kernel boundaries (for meshes too
 * small to fill the GPU, where a launch costs more than its pass); other schemes keep the launch
 * sequence.  Results are the same bits either way.  swe_tune and swe_batch_tune time both forms
 * on the resident meshes and keep the faster.  A barrier wait is bounded: should one give up,
 * the instances that were stepping fail with SWE_EHIP.  The bound is 2^21 polls, or the value of
 * the environment variable SWE_BARRIER_SPINS where set (0 forces every barrier to give up: the test
 * of that path).  SWE_EINVAL for an unknown form or a device
 * without cooperative launches.  No reference counterpart (a launch choice). */
typedef enum swe_step_form { SWE_STEP_LAUNCHES = 0, SWE_STEP_PERSISTENT = 1 } swe_step_form;
int swe_set_step_form(swe_handle *h, int form);
int swe_get_step_form(swe_handle *h, int *form);

/* -------------------------------------------------------------------------------------------
 * Batches: many meshes of different sizes resident in one device arena and advanced together,
 * each with its own parameters, boundary program, dt (its own CFL reduction), time and stopping
 * rule.  The reference runs one mesh per process; a batch runs its time loop (SOURCES/main.f90:
 * 61-136) for every instance at once, and each instance's results are bitwise those of the same
 * mesh run alone.  All instances of a batch share the scheme: viscosity_type, viscous_type,
 * alpha_limit and roundoff_fix (type_test, and with it the source term, may differ).
 * ------------------------------------------------------------------------------------------- */
typedef struct swe_batch swe_batch;

typedef struct swe_batch_caps {
    long long max_nodes;      /* node rows (each mesh takes its np rounded up to 64) */
    long long max_entries;    /* entry slots (a mesh takes nnz plus the padding of its 64-row slices, up to ~1.3 nnz) */
    long long max_boundary;   /* Dirichlet list entries (h, ux and uy lists together) */
    int max_instances;        /* meshes resident at once */
} swe_batch_caps;

typedef enum swe_instance_state {
    SWE_INSTANCE_RUNNING = 1, /* takes steps while time < tfinal and steps < max_steps */
    SWE_INSTANCE_DONE = 2,    /* reached tfinal or max_steps; its state is final */
    SWE_INSTANCE_FAILED = 3   /* a check tripped (see swe_batch_status) */
} swe_instance_state;

/* Device bytes a batch with these capacities allocates. */
unsigned long long swe_batch_bytes(const swe_batch_caps *caps);
int swe_batch_create(const swe_batch_caps *caps, swe_batch **out);
int swe_batch_destroy(swe_batch *b);

/* Admit a mesh with state un(3,np) at time t0.  It steps while time < tfinal (the loop condition
 * of SOURCES/main.f90:61) and at most max_steps steps.  Returns SWE_EFULL when the batch has no
 * room now.  The instance joins the steps queued after this call. */
int swe_batch_add(swe_batch *b, const swe_mesh_desc *mesh, const swe_params *prm, const double *un,
                  double t0, double tfinal, long long max_steps, int *id);
/* swe_batch_add for a mesh assembled by swe_assemble_device: the layout is built on the GPU from the
 * assembly (as swe_create_from_assembly does) and filled on the batch's copy stream while queued steps
 * run; bath, the Dirichlet lists (1-based) and un(3,np) come from the host.  The assembly may be
 * destroyed when the call returns.  Results are bitwise those of swe_batch_add on the same arrays. */
int swe_batch_add_assembled(swe_batch *b, const swe_assembly *a, const double *bath, int n_h_dir, const int *h_dir,
                            int n_ux_dir, const int *ux_dir, int n_uy_dir, const int *uy_dir,
                            const swe_params *prm, const double *un, double t0, double tfinal,
                            long long max_steps, int *id);
int swe_batch_set_bc(swe_batch *b, int id, int kind, double amp, double period, double level);
int swe_batch_set_bc_values(swe_batch *b, int id, const double *h_vals, const double *ux_vals,
                            const double *uy_vals);
/* Free an instance's rows (after reading its state). */
int swe_batch_remove(swe_batch *b, int id);

/* Queue nsteps steps of every running instance; returns without waiting. */
int swe_batch_step(swe_batch *b, int nsteps);
int swe_batch_sync(swe_batch *b);

/* State, time and step count of an instance (waits for queued work).  For a failed instance
 * *error is SWE_ENEG_H or SWE_ELAMBDA, else SWE_OK. */
int swe_batch_status(swe_batch *b, int id, int *state, double *time, long long *steps, int *error);
/* Instances that are done or failed, up to cap of them (waits for queued work). */
int swe_batch_poll(swe_batch *b, int *ids, int cap, int *n);
int swe_batch_get_state(swe_batch *b, int id, double *un);
/* swe_errors_create for instance `id` of the batch (compute_errors on its state, with its parameters
 * and the time passed to swe_compute_errors); the evaluator reports SWE_EINVAL once the instance has
 * been removed. */
int swe_batch_errors_create(swe_batch *b, int id, int me, const int *jj, const double *rr, swe_errors **out);

/* Pipelined serving: the host never waits for the step it just queued.
 * swe_batch_step_async queues nsteps steps, then a copy of the instance table, and returns that
 * copy's ticket.  swe_batch_finished waits for the copy of one ticket only (steps queued after it
 * keep running) and lists the instances it shows done or failed; they take no later launches.
 * swe_batch_result then reads such an instance's final state, time, step count and error on the
 * copy stream, again without waiting for the queued steps.  An instance admitted after a ticket
 * was taken is not reported by that ticket. */
int swe_batch_step_async(swe_batch *b, int nsteps, long long *ticket);
int swe_batch_finished(swe_batch *b, long long ticket, int *ids, int cap, int *n);
int swe_batch_result(swe_batch *b, int id, double *un, int *state, double *time, long long *steps,
                     int *error);

/* Free capacity: largest free run of node rows and entry slots, free instance records. */
int swe_batch_free(swe_batch *b, long long *rows, long long *entries, int *instances);

int swe_batch_tune(swe_batch *b, int steps);
int swe_batch_use_graph(swe_batch *b, int on);
int swe_batch_set_step_form(swe_batch *b, int form);   /* as swe_set_step_form */
int swe_batch_get_step_form(swe_batch *b, int *form);
void *swe_batch_stream(swe_batch *b);
int swe_batch_enable_kernel_timing(swe_batch *b, const char *kernel);
int swe_batch_kernel_timing(swe_batch *b, const char *kernel, int *launches, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* SWE_AMD_H */
