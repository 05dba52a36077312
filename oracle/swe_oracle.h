/* swe_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference explicit shallow-water step (SOURCES/update.f90, fct.f90,
 * condlim.f90, main.f90 of t-bltg/shallow-water-eric) and of its P1 set-up (SFEMaNS gauss points,
 * st_csr, qs_00_M).  It is the checker for the HIP library and the CPU baseline of the benchmark;
 * the product never links, loads or calls it.
 *
 * Conventions: 0-based CSR (row i holds entries ia[i]..ia[i+1]-1, columns ascending, diagonal
 * included, as st_csr builds them); state u is the reference's un(nc,np) in Fortran order, i.e.
 * u[nc*i+k] for node i and component k (h, qx, qy; then h*eta, h*w for the SGN model of test 13,
 * nc = 5, as main.f90:22-26 sets syst_size); c_ij is interleaved c[2*p+d].
 * Every floating-point expression keeps the reference's left-to-right evaluation order and is
 * compiled without FMA contraction.
 */
#ifndef SWE_ORACLE_H
#define SWE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_GALERKIN = 0, ORC_VISCOUS = 1, ORC_ENTROPY_VISC = 2, ORC_FCT = 3 };
enum { ORC_OK = 0, ORC_ENEG_H = 1, ORC_ELAMBDA = 2, ORC_EBADOPT = 3 };

typedef struct orc_params {
    double gravity, htiny, cfl, max_water_h, mannings;
    int type_test;        /* reference test number: selects the source term (9/10 Manning,
                             11 linear friction, 12 Coriolis; anything else none) */
    int viscosity_type;   /* ORC_GALERKIN .. ORC_FCT */
    int viscous_type;     /* 1 = 'type1', 2 = 'type2' */
    int alpha_limit;      /* inputs%if_alpha_limit */
    int roundoff_fix;     /* if_roundoff_fix (SOURCES/update.f90:15 hard-codes .TRUE.) */
    double lambda_sgn;    /* inputs%lambdaSGN (SOURCES/read_data.f90:92-95), test 13 only */
} orc_params;

typedef struct orc_ctx orc_ctx;

/* set-up */
long orc_csr_nnz(int np, int me, const int *jj);
int orc_assemble(int np, int me, const int *jj, const double *rr,
                 int *ia, int *ja, int *diag, double *mass, double *lumped, double *cij);

/* solver context: copies the arrays it is given */
orc_ctx *orc_create(int np, int nnz, const int *ia, const int *ja, const double *cij,
                    const double *lumped, const double *bath, const orc_params *prm);
void orc_destroy(orc_ctx *c);
void orc_set_params(orc_ctx *c, const orc_params *prm);
/* number of components of the context's states (3, or 5 for type_test 13) */
int orc_syst_size(const orc_ctx *c);
/* node coordinates rr(2,np), needed by the bc_kind 3 boundary program of orc_run */
void orc_set_coords(orc_ctx *c, const double *rr);

/* pointwise pieces */
double orc_lambda_vacc(const double *ul, const double *ur, const double *nij,
                       double gravity, double htiny, int *err);
void orc_lambda_batch(int n, const double *rec8, double gravity, double htiny, double *out, int *err);
void orc_fct_positivity(int np, const int *ia, const int *ja, const double *f,
                        const double *unext_scaled, double htiny, double *lij);
void orc_fct_generic(int np, const int *ia, const int *ja, const double *f, const double *maxn,
                     const double *minn, const double *dg, double *unext);

/* step pieces on a context */
int orc_compute_dij(orc_ctx *c, const double *un, double *dij_out);
double orc_compute_dt(orc_ctx *c, const double *un, int *err);
int orc_euler(orc_ctx *c, const double *un, double *unext, double dt, double *dij_out);

/* Fixed-step SSP-RK3 loop of SOURCES/main.f90:61-92.  bc_kind 0: no Dirichlet lists allowed;
 * 1: walls (qx = 0 on nodes_ux, qy = 0 on nodes_uy); 2: walls plus the inflow
 * h = max(level - z + amp*orc_sin_rn(2*pi*t/period), 0) on nodes_h; 3: the reference's bdy
 * (main.f90:176-188), sol_anal of the context's type_test at the listed nodes (orc_set_coords
 * first).  dts receives dt per step. */
int orc_run(orc_ctx *c, double *un, int nsteps, double *time_io, double *dts,
            int bc_kind, double amp, double period, double level,
            int nh, const int *nodes_h, int nux, const int *nodes_ux, int nuy, const int *nodes_uy);

/* sin(x) rounded to nearest (double-double evaluation, one final rounding), |x| < 2^29. */
double orc_sin_rn(double x);

/* sol_anal(k, (x, y), t) (SOURCES/condlim.f90:311-831) for the analytic tests 3 to 7, 11, 12 and
 * 13, components k = 1 (h), 2 (qx), 3 (qy) (and for test 13 k = 4 (h*eta), 5 (h*w) at t <= 1e-10,
 * the only times the reference defines them); *ok = 0 for other cases. */
double orc_sol_anal(int type_test, int k, double x, double y, double t, double gravity, double htiny, int *ok);

/* The quadrature norms of compute_errors (SOURCES/main.f90:198-263, tests 3-7, 11, 12 and 13; u holds
 * 5 components per node for test 13, of which h, qx, qy enter): r_gauss (303-322),
 * sol_anal at the Gauss points, ns_anal_l1 and ns_anal_0 (SFEMaNS/FEMSUB_OBJET/fem_tn.f90:
 * 1679-1716, 1513-1553) of u(k,:) - exact and of 0 - exact, sequential sums in element and Gauss
 * point order.  jj is 0-based jj[3*m+n]; out[4*(k-1) + {0,1,2,3}] = L1 error, L1 norm, L2 error,
 * L2 norm.  terms (may be NULL) receives every summand: terms[(4*(k-1)+q)*3*me + 3*m + l]
 * (before the square root of the L2 sums).  Returns ORC_EBADOPT for tests without sol_anal here. */
int orc_compute_errors(int np, int me, const int *jj, const double *rr, const double *u, int type_test,
                       double t, double gravity, double htiny, double *out, double *terms);

/* The diagnostic the time loop writes to unit 10 for tests 1, 2 and 10 (SOURCES/main.f90:94-99):
 * ns_l1 (SFEMaNS/FEMSUB_OBJET/fem_tn.f90:2119-2153), the sum over elements and Gauss points of
 * |sum_n f(jj(n,m)) ww(n,l)| rj(l,m), of f = qx, qy and h*sqrt(h), sequential in element and Gauss
 * point order; out = {q1, q2, q3, 0.5*(q1+q2)/(gravity*q3)}.  u holds nc components per node (h, qx,
 * qy first).  terms (may be NULL) receives every summand: terms[f*3*me + 3*m + l]. */
void orc_ns_l1_ratio(int np, int me, const int *jj, const double *rr, const double *u, int nc, double gravity,
                     double *out, double *terms);

#ifdef __cplusplus
}
#endif
#endif
