/*
 * wql_oracle.h -- CPU restatement (FP64) of the WaveQLab3D-Q per-block time
 * step.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Array layout (all arrays are plain C arrays, no ghost nodes):
 *   node (x,y,z), 1-based as in the reference, x = q fastest:
 *       idx = (x-1) + nx*((y-1) + ny*(z-1)),   N = nx*ny*nz
 *   component-major: F[c*N + idx], c = 0..8 (vx,vy,vz,sxx,syy,szz,sxy,sxz,syz)
 *   metrics: metx[m*N + idx] = (q_x, r_x, s_x)[m], mety = (q_y, r_y, s_y),
 *            metz = (q_z, r_z, s_z)     (reference grid.f90:398-471)
 *   material M[k*N + idx] = (lambda, mu, rho)[k]   (material.f90:696-702)
 *   memory variables eta[(c6*nmech + i)*N + idx], c6 = 0..5 -> components 4..9
 *   PML auxiliaries: full-block [9][N] arrays per face (only slab nodes used)
 *   face frames: nl/nm/nn[f][k*NF + fidx], f = 0..5 = Lx,Rx,Ly,Ry,Lz,Rz,
 *     q faces fidx = (y-1)+ny*(z-1); r faces (x-1)+nx*(z-1); s faces (x-1)+nx*(y-1)
 */
#ifndef WQL_ORACLE_H
#define WQL_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_SRC_GAUSSIAN = 0, ORC_SRC_COSBELL = 1, ORC_SRC_RICKER = 2 };

typedef struct orc_source {
  int type;
  int order;          /* 2 or 4 (singular_source); 0 = mollified (set_moment_tensor_smooth) */
  int alpha;          /* 0, 25, 50, 100 for order 4 */
  int near[3];        /* nearest node, 1-based */
  double duration, t_init;
  double mXX, mYY, mZZ, mXY, mXZ, mYZ;
  double loc[3];
  double near_phys[3];
} orc_source;

typedef struct orc_block {
  int nx, ny, nz;
  int curvilinear;    /* type_of_mesh: 0 cartesian, 1 curvilinear */
  int nmech;          /* 0 elastic, 4 = anelastic-Q4, 8 = anelastic-Q8 */
  int nsrc;
  int bc[6];          /* Lx,Rx,Ly,Ry,Lz,Rz; 1 char., 2 free surface, 3 clamped, <=0 none */
  int pml[6];
  int npml[6];
  /* SAT form: 0 = BCm/BCp + structure as in BoundaryConditions.f90:783-924
   * and RHS_Interior.f90:1584-1768; 1 = the alpha/theta form kept in
   * comments at BoundaryConditions.f90:833-841 / 905-913 (velocity penalty
   * alpha*norm/rho*F with alpha = 2, 1, 0 for codes 1, 2, 3; stress penalty
   * theta*norm*(...) with theta = 0, 0, -1; no structure), which produced
   * the archived golden files test_problems/truth/1; no PML faces */
  int sat_form;
  double hq, hr, hs, tau0;
  double tau[8], weight[8];
  double *F, *DF;
  double *metx, *mety, *metz, *J;
  double *M;
  double *eta, *Deta;
  double *Qs_inv, *Qp_inv;
  double *pQ[6], *pDQ[6];
  double *nl[6], *nm[6], *nn[6];
  orc_source *src;
  double *X;          /* node coordinates [3][N] (x, y, z); read by mollified sources */
} orc_block;

/* KC(5,4) low-storage coefficients (time_step.f90:53-62): A[5], B[5], C[6] */
void orc_rk_coefficients(double *A, double *B, double *C);

/* SBP-6 first-derivative closure table (JU_xJU_yJU_z6.f90:74-182), 6x9 */
void orc_sbp_closure(double *table54);

/* the SBP-6 first-derivative rows of JJU_x6 / D6x applied to u[n] */
void orc_apply_d1(int n, double h, const double *u, double *du);

/* derivatives() (grid.f90:398-471): metrics and J from coordinates X[3][N] */
void orc_metrics(int nx, int ny, int nz, const double *X,
                 double *metx, double *mety, double *metz, double *J);

/* local_orth_vectors_{q,r,s} (unit_normals.f90:8-168) for all six faces */
void orc_normals(orc_block *b);

/* fixed-coefficient Q4/Q8 tau/weights (anelastic_q{4,8}_model.f90) */
void orc_q_coefficients(int nmech, double fmin, double fmax, double *tau, double *weight);

/* per-node unrelaxed-modulus correction (material.f90:425-453 / 258-280);
 * returns 0 or -1 if weight/Q sum >= 1 at some node */
int orc_q_modulus_correction(orc_block *b, double fref);

/* stage pieces, in reference order */
void orc_scale_rates(orc_block *b, double A);               /* fields.f90:51-155 */
void orc_set_rates(orc_block *b);                           /* elastic.f90:13-29 */
void orc_set_moment_tensor(orc_block *b, double t);         /* moment_tensor.f90:875-935, 1255-1330 */
void orc_impose_boundary_condition(orc_block *b);           /* RHS_Interior.f90:1584-1768 */
void orc_update_fields(orc_block *b, double inc);           /* fields.f90:158-263 */

/* time_step_RK (time_step.f90:68-150) for nsteps steps; t is in/out.
 * If nrec > 0, rec_idx[nrec] (0-based flat node index) are sampled at stage 1
 * of each step (time t0) into seis[(step*nrec + r)*9 + c]. */
void orc_time_step(orc_block *b, double dt, int nsteps, double *t,
                   int nrec, const long *rec_idx, double *seis);

/* OpenMP threads the node loops use (1 without OpenMP) */
int orc_num_threads(void);
void orc_set_num_threads(int n);

/* source_time (moment_tensor.f90:813-872) + Ricker (new) */
double orc_source_time(int type, double t, double t0, double w);

/* ---- two-block q-interface (wql_oracle_iface.c) ------------------------ */
enum { ORC_IFACE_LOCKED = 0, ORC_IFACE_SLIP_WEAKENING_TPV5 = 1, ORC_IFACE_SLIP_WEAKENING_TPV28 = 2,
       /* coupling 'rate-and-state_friction' with problem 'TPV101' / 'TPV102'
        * (the values of the library's WQL_IFACE_* codes; 3 is its caller-table
        * slip weakening, which the oracle has no use for) */
       ORC_IFACE_RATE_STATE_TPV101 = 4, ORC_IFACE_RATE_STATE_TPV102 = 5,
       /* coupling 'linear_friction' (alpha from orc_iface.alpha) and
        * 'rate-and-state_friction' with problem 'rate-weakening' */
       ORC_IFACE_LINEAR_FRICTION = 6, ORC_IFACE_RATE_STATE_RW = 7 };

/* rate-and-state pieces (Interface_Condition.f90), exposed for their pins:
 * Regula_Falsi (:1170-1239) on f (:1241-1265) from the guess V, returning the
 * slip rate; and the state rate of TPV101 / TPV102 (:431-436) */
double orc_rs_regula_falsi(double V, double Phi, double eta, double sigma_n, double psi,
                           double V0, double a);
double orc_rs_state_rate(double W, double Vel, double L0, double f0, double b, double V0);
/* the friction parameter a of TPV101 / TPV102 at (y, z) (:386-400, with
 * smooth_boxcar :2460-2476) */
double orc_rs_a(double y, double z);
/* problem 'rate-weakening': a(y, z) and the flash-weakening velocity
 * V_w(y, z) (:410-416, 438-444), and the state rate
 * -V/L0 (a asinh(V/(2 V0) e^(W/a)) - Fss) (:438-444) */
double orc_rw_a(double y, double z);
double orc_rw_vw(double y, double z);
double orc_rw_state_rate(double W, double Vel, double a, double Vw);
/* its prestress (n, m, l) on a face with frame n, m, l at (y, z), time t
 * (:2387-2426) */
void orc_rw_prestress(const double *n, const double *m, const double *l, double y, double z,
                      double t, double T0[3]);
/* hat variables P (minus side), Q (plus side) in (n, m, l) from the rotated
 * grid values U, V (velocities 0..2, tractions 3..5) of Welded_Interface
 * (:515-590) and of Linear_Friction (:639-742) with tau = alpha v */
void orc_welded_hats(const double U[6], const double V[6], double Z1s, double Z2s, double Z1p,
                     double Z2p, double P[6], double Q[6]);
void orc_linear_friction_hats(const double U[6], const double V[6], double Z1s, double Z2s,
                              double Z1p, double Z2p, double alpha, double P[6], double Q[6]);

/* Interface between the x-high face of block L and the x-low face of block
 * R (same ny, nz).  Face arrays [k][NF], fidx = (y-1) + ny*(z-1).
 *   S, DS: slip (3 local components n, m, l) and slip path length, and their
 *   rates (iface.f90:36-44: S = 0, DS = 1e40 initially); W, DW: state;
 *   yf, zf: y and z coordinates of the face nodes (prestress, strength);
 *   stage-1 outputs Uhat, Vhat [6] (slip velocity (local), hat tractions +
 *   prestress), Svel [3], trup (1e9 initially). */
typedef struct orc_iface {
  int ny, nz;
  int coupling;
  /* TPV28 initial stress (xx, yy, zz, xy, xz, yz): the source's values are
   * (-60, 0, 60, 0, 29.38, 0), initial_stress_condition.f90:78-84; the
   * archived golden files test_problems/truth/2 were made with zz = -60 */
  double s0[6];
  double *yf, *zf;
  double *S, *DS, *W, *DW;
  double *Uhat, *Vhat, *Svel, *trup;
  double alpha;     /* linear friction: tau = alpha v */
} orc_iface;

/* enforce_iface_conditions for one side: side 2 = block L's call (SAT into
 * L), side 1 = block R's call (SAT into R); both add to DS */
/* rate-and-state: I->W is the state psi (iface.f90:38, 0 there; or the
 * caller's initial state), I->DW its rate; the evaluation of block L's side
 * adds the state rate once per stage, like DS */
void orc_iface_enforce(orc_iface *I, orc_block *L, orc_block *R, double t, int stage, int side);
void orc_iface_scale(orc_iface *I, double A);     /* scale_rates_iface (iface.f90:63-73) */
void orc_iface_update(orc_iface *I, double inc);  /* update_fields_iface (iface.f90:76-88) */

#ifdef __cplusplus
}
#endif
#endif
