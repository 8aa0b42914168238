/*
 * wql3d.h -- C ABI of libwql3d_hip.so, the MI355X (gfx950) implementation of
 * the WaveQLab3D-Q per-block RK time step: 6th-order SBP traditional RHS
 * (Cartesian and curvilinear), anelastic Q4/Q8 memory variables, PML
 * auxiliary fields, SAT boundary penalties, moment-tensor source injection
 * and the Kennedy-Carpenter (5,4) low-storage RK update.
 *
 * One block of one rank lives on one GPU.  There is no halo exchange: on one
 * rank the reference's exchanges are sends to MPI_PROC_NULL
 * (mpi3dcomm.f90:478-527), so ghost nodes never carry data.
 *
 * Conventions
 *  - Every entry point returns 0 on success or a WQL_E* code; the message of
 *    the last failure on the calling thread is wql_last_error().  A Fortran
 *    shim maps a non-zero code to fatal_local('RUN-HIP-<code>', message)
 *    (diagnostics.f90:175-204).
 *  - Host arrays are Fortran arrays passed by c_loc of their first element
 *    together with their bounds (wql_box), so arrays with ghost nodes and
 *    non-unit lower bounds are read and written in place.  If
 *    desc->arrays_on_device is 1 the same pointers are device pointers.
 *  - Calls are stream-ordered on the block's own HIP stream; download and
 *    sample synchronise.  No entry point allocates device memory except
 *    wql_block_create / upload / download / set_* (never inside a step).
 *  - Calls on one block come from one host thread.
 */
#ifndef WQL3D_H
#define WQL3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
enum {
  WQL_OK = 0,
  WQL_EINVAL = 1,   /* bad argument / descriptor */
  WQL_EHIP = 2,     /* HIP runtime error */
  WQL_ENOMEM = 3,   /* device memory capacity exceeded (RUN-GPU-MEM-002) */
  WQL_ESTATE = 4,   /* call not valid in the block's current state */
  WQL_ENODEV = 5    /* no GPU visible */
};

enum { WQL_CARTESIAN = 0, WQL_CURVILINEAR = 1 };       /* type_of_mesh */
/* arithmetic mode: FP64; FP32 stencil/metric/memory arithmetic with the RK
 * state F, DF in FP64 (FP32_MIXED; the block also keeps F rounded to float as
 * the stencils' operands, 72 B per node); or all FP32 with a
 * Kahan-compensated F update (FP32_KAHAN, fused stage only) */
enum { WQL_FP64 = 0, WQL_FP32_MIXED = 1, WQL_FP32_KAHAN = 2 };
enum { WQL_BC_NONE = 0, WQL_BC_CHARACTERISTIC = 1,     /* BoundaryConditions.f90:28-30 */
       WQL_BC_FREE_SURFACE = 2, WQL_BC_CLAMPED = 3 };
enum { WQL_SRC_GAUSSIAN = 0, WQL_SRC_COSBELL = 1, WQL_SRC_RICKER = 2 };
/* SAT penalty form of the boundary faces: BCm/BCp followed by structure()
 * (BoundaryConditions.f90:783-924, RHS_Interior.f90:1584-1768), or the
 * alpha/theta form kept in comments at BoundaryConditions.f90:833-841 /
 * 905-913 (velocity penalty alpha*norm/rho*F, alpha = 2, 1, 0 for codes 1, 2,
 * 3; stress penalty theta*norm*(...), theta = 0, 0, -1; no structure), the
 * form that produced the archived golden files test_problems/truth/1.  The
 * alpha/theta form takes no PML faces. */
enum { WQL_SAT_BC = 0, WQL_SAT_ALPHA_THETA = 1 };

/* upload/download masks */
enum {
  WQL_F = 1u, WQL_DF = 2u, WQL_GEOMETRY = 4u, WQL_MATERIAL = 8u,
  WQL_MEMORY = 16u,   /* eta, Deta */
  WQL_QINV = 32u,     /* Qs_inv, Qp_inv */
  WQL_PML = 64u,      /* PML Q, DQ */
  WQL_FACES = 128u,   /* n_l, n_m, n_n */
  WQL_ALL = 255u
};

/* Fortran bounds of the first three dimensions of an array
 * (for face arrays only lo[0..1], hi[0..1] are used) */
typedef struct wql_box {
  int32_t lo[3];
  int32_t hi[3];
} wql_box;

/* Block descriptor: the fields of block_type / block_fields / block_grid_t /
 * block_material / block_pml / block_boundary (datatypes.f90:24-247) that the
 * time step reads or writes. */
typedef struct wql_block_desc {
  int32_t nq, nr, ns;        /* nodes of the block, nodes are 1..n per axis */
  int32_t mesh;              /* WQL_CARTESIAN / WQL_CURVILINEAR */
  int32_t precision;         /* WQL_FP64 / WQL_FP32_MIXED / WQL_FP32_KAHAN */
  int32_t nmech;             /* 0 elastic, 4 anelastic-Q4, 8 anelastic-Q8 */
  int32_t bc[6];             /* Lx,Rx,Ly,Ry,Lz,Rz (boundary_type Lx..Rz) */
  int32_t pml[6];            /* PMLB(1..6)%pml */
  int32_t npml[6];           /* PMLB(1..6)%N_pml */
  int32_t arrays_on_device;  /* 0: host pointers, 1: device pointers */
  double hq, hr, hs;         /* G%hq, G%hr, G%hs = 1/(n-1) */
  double tau0;               /* B%tau0 (block.f90:242) */
  double tau[8], weight[8];  /* tau_Q / weight_Q (Q4) or tau_Q8 / weight_Q8 (Q8) */
  wql_box body;              /* bounds of F, DF, metric*, J, M, eta*, Q*_inv */
  const double *metricx, *metricy, *metricz;  /* body x 3 */
  const double *J;                            /* body */
  const double *M;                            /* body x 3: lambda, mu, rho */
  double *F, *DF;                             /* body x 9 */
  double *eta[6], *Deta[6];                   /* eta4..eta9 (+Q/Q8 suffix): body x nmech */
  const double *Qs_inv, *Qp_inv;              /* body */
  wql_box pml_box[6];                         /* bounds of PMLB(f)%Q / %DQ */
  double *pml_Q[6], *pml_DQ[6];               /* pml_box x 9 */
  wql_box face_box[6];                        /* bounds of B(f)%n_l/n_m/n_n (2-D) */
  const double *n_l[6], *n_m[6], *n_n[6];     /* face_box x 3; needed on faces with
                                                 bc > 0 and on interface faces */
  int32_t sat_form;                           /* WQL_SAT_BC / WQL_SAT_ALPHA_THETA */
} wql_block_desc;

/* moment tensor point source (moment_tensor.f90: moment_tensor type) */
typedef struct wql_source {
  int32_t type;              /* WQL_SRC_* ; for RICKER duration = peak frequency */
  int32_t order;             /* 2 or 4 (singular_source); 0 = mollified:
                                set_moment_tensor_smooth (moment_tensor.f90:
                                1255-1330, the reference's mollify_source),
                                needs wql_block_set_coordinates first */
  int32_t alpha;             /* 0, 25, 50, 100 (order 4 stencil) */
  int32_t near[3];           /* nearest node, 1-based */
  double duration, t_init;
  double mXX, mYY, mZZ, mXY, mXZ, mYZ;
  double location[3];
  double near_phys[3];
} wql_source;

typedef struct wql_block wql_block;

/* ---- accelerator_runtime seam (accelerator_runtime_cpu.f90:19-64) ------- */
/* initialize_accelerator_runtime: bind rank to device local_rank % ndev */
int wql_runtime_init(int32_t local_rank, int32_t *device);
/* finalize_accelerator_runtime */
int wql_runtime_finalize(void);
/* enforce_device_memory_capacity (domain.f90:405): WQL_ENOMEM if the bytes
 * exceed free device memory; env WQL3D_GPU_MEMORY_BYTES overrides capacity */
int wql_capacity_check(int64_t predicted_bytes);

/* ---- device_data seam (device_data_cpu.f90:8-26) ------------------------- */
/* enter_domain_device_data: create device mirrors (hipMalloc, once) */
int wql_block_create(const wql_block_desc *desc, wql_block **out);
/* device bytes held by the block (persistent_memory.f90:47-105 analogue) */
int64_t wql_block_device_bytes(const wql_block *b);
/* copy host arrays selected by mask to the device mirrors */
int wql_block_upload(wql_block *b, uint32_t mask);
/* update_domain_host_data: copy device mirrors back to the host arrays */
int wql_block_download(wql_block *b, uint32_t mask);
/* set F=0, DF=1e40, eta=Deta=0, PML Q=0, DQ=1e40 on the device, as
 * init_fields / init_anelastic_Q*_properties / init_pml do */
int wql_block_init_state(wql_block *b);
/* exit_domain_device_data */
void wql_block_destroy(wql_block *b);

/* ---- sources and receivers ---------------------------------------------- */
/* init_moment_tensor result (moment_tensor.f90:8-242); n <= 8 */
int wql_block_set_sources(wql_block *b, int32_t n, const wql_source *src);
/* node coordinates B%G%X (datatypes.f90 block_grid_t), read by mollified
 * sources: X[c*nx*ny*nz + idx], c = x, y, z, idx = x + nx*(y + ny*z) (0-based),
 * a host array, or a device array when on_device != 0 */
int wql_block_set_coordinates(wql_block *b, const double *X, int32_t on_device);
/* receivers at nodes ijk[3*n] (1-based), sampled at stage 1 of each step
 * (time_step.f90:137) into a device buffer of nslots steps */
int wql_block_set_receivers(wql_block *b, int32_t n, const int32_t *ijk, int32_t nslots);
/* copy recorded samples [slot][receiver][9] for slots 0..nslots-1 to host */
int wql_block_get_seismograms(wql_block *b, int32_t nslots, double *out);

/* ---- rk_vector_kernels seam (rk_vector_cpu.f90:7-19) -------------------- */
/* scale_rate_array over DF, Deta*, PML DQ of the block (fields.f90:51-155) */
int wql_rk_scale(wql_block *b, double A);
/* update_state_array over F, eta*, PML Q (fields.f90:158-263) */
int wql_rk_update(wql_block *b, double Bdt);
/* the same two kernels on any device array of n doubles (bit-exact with the
 * CPU expressions rate = c*rate and state = state + inc*rate) */
int wql_scale_rate_array(double *rate, int64_t n, double coefficient);
int wql_update_state_array(double *state, const double *rate, int64_t n, double increment);

/* ---- device mirrors of single host arrays: the present table that the
 * reference's OpenACC backend keeps (acc enter data copyin / acc_is_present /
 * update device / update self / exit data; rk_vector_openacc.f90:8-55,
 * tests/rk_vector_test_data_openacc.f90:5-12), for the per-array rk_vector
 * seam.  A host range is present when it lies inside one entered array. */
int wql_array_enter(const void *host, int64_t bytes);      /* device mirror + copy in */
int wql_array_exit(const void *host);                      /* drop the mirror (no copy out) */
int wql_array_present(const void *host, int64_t bytes);    /* 1 present, 0 not */
int wql_array_update_device(const void *host, int64_t bytes);
int wql_array_update_self(void *host, int64_t bytes);
/* scale_rate_array / update_state_array (rk_vector_cpu.f90:7-19) on one
 * present host array of n doubles (the whole array, ghost nodes included),
 * as the OpenACC backend runs them (rk_vector_openacc.f90:8-55): update the
 * mirror from the host, the kernel on the mirror (bit-exact with the CPU
 * expressions), the result back to the host.  WQL_ESTATE when an array is
 * not present (the OpenACC backend's 'not present on device' stop). */
int wql_scale_rate_present(double *rate, int64_t n, double coefficient);
int wql_update_state_present(double *state, const double *rate, int64_t n, double increment);
/* the block created with host array desc->F == host_F (the swap-in seams that
 * receive a block_type find its device mirror through F%F's address) */
int wql_block_find(const double *host_F, wql_block **out);

/* ---- set_rates_elastic seam (elastic.f90:13-29): RHS_Center +
 *      RHS_near_boundaries incl. Q4/Q8 and PML: DF += L(F) ------------------ */
int wql_rhs(wql_block *b);
/* try_traditional_cartesian_rhs seam (traditional_cartesian_rhs_openacc.f90:10-123):
 * strict-interior only; *handled = 0 when not eligible (then nothing ran) */
int wql_rhs_interior(wql_block *b, int32_t *handled);
/* set_moment_tensor (moment_tensor.f90:875-935) at time t */
int wql_set_moment_tensor(wql_block *b, double t);
/* impose_boundary_condition (RHS_Interior.f90:1584-1768) */
int wql_sat(wql_block *b);

/* ---- fused paths --------------------------------------------------------- */
/* one RK stage: DF = A*DF + L(F) + source(t) - SAT; F += Bdt*DF (F ping-pong),
 * memory variables and PML auxiliaries likewise */
int wql_stage_fused(wql_block *b, double A, double Bdt, double t);
/* time_step_RK (time_step.f90:68-150): nsteps steps of 5 fused stages;
 * *t advanced as the reference advances D%t; receivers recorded */
int wql_time_step(wql_block *b, double dt, int32_t nsteps, double *t);

/* ---- host metric pipeline (grid.f90:398-471 derivatives(), metrics.f90) on
 * the device: X[3] coordinates (node-major, x fastest, no ghosts, device
 * pointer) -> metricx/y/z[3] and J (same layout, device pointers) -------- */
int wql_grid_metrics(int32_t nq, int32_t nr, int32_t ns, const double *X,
                     double *metricx, double *metricy, double *metricz, double *J);

/* ---- diagnostics --------------------------------------------------------- */
const char *wql_last_error(void);
/* close_domain's shutdown check (domain.f90:520-545): out[0] = max|F| over
 * the block's nine field components, out[1] = max|eta| over its memory
 * variables (0 for an elastic block), out[2] = the number of non-finite
 * values among them (the reference stops with RUN-ELASTIC-001 / RUN-Q4-001 /
 * RUN-Q8-001 when there is one).  out is a host array of 3; the call
 * allocates its partial results on the device for the call and synchronises
 * the block's stream. */
int wql_block_max_state(wql_block *b, double *out);
/* No counterpart in the reference: wql_block_max_state for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The same kernels, with their partial
 * results in a fixed 48 KiB device array of the library (2048 partials; one
 * copy per GPU, held until the process ends; the first pass never uses more
 * partials, whatever the block size) instead of an allocation; the three
 * results go to out_dev, a DEVICE array of 3 doubles, stream-ordered on the
 * block's stream (no synchronisation).  Calls that overlap on one GPU share
 * that array and race. */
int wql_block_max_state_noalloc(wql_block *b, double *out_dev);
/* synchronise the block's stream */
int wql_block_sync(wql_block *b);
/* the block's HIP stream (hipStream_t) for event timing by the caller */
void *wql_block_stream(wql_block *b);
/* Kernel timing of fused stages (the reference times each step with
 * cpu_time around time_step_RK, main.f90:120-132; here HIP events on the
 * block's stream time the kernels of each stage).  enable != 0 records, for
 * every following wql_stage_fused / wql_time_step stage, the durations of
 * ms[0] the strict-interior kernel, ms[1] the node kernel over the rest of
 * the block and ms[2] the memory-variable kernel (0 when a part does not
 * run); wql_block_kernel_times synchronises, returns the sums in
 * milliseconds and the number of stages recorded since the last call, and
 * starts a new count. */
int wql_block_set_timing(wql_block *b, int32_t enable);
int wql_block_kernel_times(wql_block *b, double ms[3], int32_t *stages);

/* ---- two-block interface (nblocks = 2, interface direction 'q') ----------
 * The x-high face (Rx, bc 0) of block L is coupled to the x-low face (Lx,
 * bc 0) of block R; both blocks have the same ny, nz and live on this GPU.
 * Per stage, as enforce_bound_iface_conditions (domain.f90:615-644) with
 * Impose_Interface_Condition (RHS_Interior.f90:1771-1843), Couple_Interface_x
 * (CouplingForcing.f90:9-463) and Interface_RHS (Interface_Condition.f90:
 * 12-513): hat variables from both faces (welded, or TPV5 linear slip-weakening
 * friction with Tau_strength / prestress TPV5, Interface_Condition.f90:
 * 1283-1299 / 1802-1830), the characteristic SAT forcing of block L (its side
 * 2) and of block R (its side 1), the slip rate DS(1:3) = Q - P and
 * DS(4) = |slip rate| (one increment per stage, as each MPI rank holds its own
 * copy of the interface), S += Bdt*DS (scale_rates_iface / update_fields_iface,
 * iface.f90:58-85).  The stage-1 fault output (write_fault_output /
 * write_hat_output, domain.f90:782-847) is recorded into device slots.
 * The interface reads FP64 copies of the coupled faces' frames and material
 * (taken at wql_iface_create and at every upload of WQL_FACES / WQL_MATERIAL
 * of a coupled block), in every precision. */
/* Slip-weakening friction is evaluated from per-node tables: the prestress
 * (T0_n, T0_m, T0_l) in each block's face frame (prestress(), Interface_Condition.f90:
 * 1802-2440) and the parameters of the linear slip-weakening strength
 * tau_str = sigma_n*(fs - (fs - fd)*min(D, Dc)/Dc) (Tau_strength, :1268-1797),
 * fs including the problem's high-strength barrier.  The TPV5 and TPV28 tables
 * are built at wql_iface_create (TPV5 :1283-1299, 1821-1830; TPV28 :1459-1475,
 * 2049-2079 with initial_stress_condition.f90:78-84); WQL_IFACE_SLIP_WEAKENING
 * takes them from the caller (e.g. filled by the reference's own prestress()). */
/* Rate-and-state friction, problems TPV101 / TPV102 (coupling
 * 'rate-and-state_friction', Interface_Condition.f90:375-436): per node the
 * slip rate V solves V + a sigma_n / eta asinh(V / (2 V0) e^(psi / a)) =
 * Phi / eta by the reference's regula falsi (:1170-1265; FP64 in every
 * precision), the hats follow rate_state_friction (:918-1067), and the
 * state psi evolves as dpsi/dt = b V0 / L0 e^(-(psi - f0) / b) - V b / L0
 * (:431-436) with L0 0.02, f0 0.6, b 0.012, V0 1e-6 and a(y, z) 0.008 inside
 * the 15 km box, 0.016 outside (smooth_boxcar); the prestress is the traction
 * of the stress (-120, 0, -120, 0, 75, 0) plus the nucleation 25 F(r) G(t)
 * on T0_l (:2324-2378).  The initial state is desc->W0 ([nz][ny]) or 0 as
 * init_iface sets it (iface.f90:38; the reference's Init_Vel_State,
 * :2722-2850, which would set the steady state, is not called there). */
/* Linear friction (coupling 'linear_friction', Interface_Condition.f90:
 * 360-366, 639-742): tau = alpha v on the shear components.  The hats are
 * the closed form of Linear_Friction with alpha = desc->alpha (the reference
 * fixes alpha = 1e14, which a negative desc->alpha selects; 0 is a
 * frictionless interface).  The reference's routine leaves its u_out(2)
 * (the minus side's m velocity hat) unset; here it is u_m_hat of the same
 * derivation.  The fault output adds desc->T0L / T0R (the problem's
 * prestress, [3][nz][ny]) when given, else zero. */
/* Rate-and-state friction, problem 'rate-weakening' (Interface_Condition.
 * 375-446, 2387-2426; initial_stress_condition.f90:184-191; Init_Vel_State
 * :2822-2845): L0 0.5144, f0 0.7, b 0.02, V0 1e-6, a(y, z) 0.016 inside the
 * 20 km box, 0.032 outside (smooth_boxcar width 2.5); the prestress is the
 * traction of the stress with psi = 50 degrees (sxx -126, sxz 42, szz from
 * the angle, syy their mean) plus a Gaussian pulse 2 tau_pulse F(r) on T0_l
 * from t > 0 (tau_pulse = -0.2429 T0_n, r0 1.35 km around (y, z) = (8, 0));
 * the state evolves as dpsi/dt = -V / L0 (a asinh(V / (2 V0) e^(psi / a))
 * - Fss) with the flash-weakened steady state Fss = 0.13 + (f0 - (b - a)
 * ln(V / V0) - 0.13) / (1 + (V / V_w)^8)^(1/8), V_w(y, z) = 0.17 inside the
 * box, 5.17 outside.  The slip-rate solve is the TPV101 one.  The initial
 * state is desc->W0 or 0 (Init_Vel_State's value is 0.4367). */
enum { WQL_IFACE_LOCKED = 0, WQL_IFACE_SLIP_WEAKENING_TPV5 = 1,
       WQL_IFACE_SLIP_WEAKENING_TPV28 = 2, WQL_IFACE_SLIP_WEAKENING = 3,
       WQL_IFACE_RATE_STATE_TPV101 = 4, WQL_IFACE_RATE_STATE_TPV102 = 5,
       WQL_IFACE_LINEAR_FRICTION = 6, WQL_IFACE_RATE_STATE_RW = 7 };
typedef struct wql_iface_desc {
  int32_t coupling;          /* WQL_IFACE_*: coupling 'locked' / 'slip-weakening_friction'
                                with problem 'TPV5' / 'TPV28' / caller tables */
  int32_t arrays_on_device;  /* 0: the arrays below are host pointers, 1: device pointers */
  const double *y, *z;       /* coordinates of the interface nodes, [nz][ny] (y fastest):
                                the y and z of block L's x-high face (G%X(pq, :, :, 2:3)) */
  const double *s0;          /* TPV28 only: initial stress (xx, yy, zz, xy, xz, yz), host
                                array of 6; null = (-60, 0, 60, 0, 29.38, 0) */
  const double *T0L, *T0R;   /* WQL_IFACE_SLIP_WEAKENING: prestress [3][nz][ny] in the
                                (n, m, l) frame of block L's face / block R's face;
                                WQL_IFACE_LINEAR_FRICTION: the output's prestress (or null) */
  const double *fs, *fd, *Dc;  /* WQL_IFACE_SLIP_WEAKENING only: [nz][ny] each */
  const double *W0;          /* WQL_IFACE_RATE_STATE_*: initial state [nz][ny]; null = 0 */
  double alpha;              /* WQL_IFACE_LINEAR_FRICTION: tau = alpha v; < 0 = 1e14 */
} wql_iface_desc;
typedef struct wql_iface wql_iface;
/* init_iface (iface.f90:13-55): S = 0, DS = 1e40, rupture time 1e9; for
 * TPV28 the blocks' face frames must be uploaded (WQL_FACES) before.  Both
 * blocks share one precision (any WQL_FP*); the interface arithmetic and its
 * slip state are FP64 in all three */
int wql_iface_create(wql_block *L, wql_block *R, const wql_iface_desc *desc, wql_iface **out);
void wql_iface_destroy(wql_iface *I);
/* one RK stage of both blocks and the interface; stage1 != 0 records the fault
 * output of time t into the next slot (if slots are set) */
int wql_iface_stage(wql_iface *I, double A, double Bdt, double t, int32_t stage1);
/* time_step_RK (time_step.f90:68-150) of the two-block domain: nsteps steps of
 * 5 stages, *t advanced as D%t; receivers of both blocks and the fault output
 * recorded at stage 1 */
int wql_iface_time_step(wql_iface *I, double dt, int32_t nsteps, double *t);
/* device slots for nslots steps of fault output; each slot holds, per node
 * of the face ([nz][ny]), 39 planes: Uface (block L face fields, 9), Vface
 * (block R face fields, 9), S (4), Svel (3), Uhat (6), Vhat (6), rupture time (1),
 * friction state psi (1; 0 for slip-weakening and locked coupling), in the order
 * and meaning of the reference's .Uface/.Vface/.S/.Svel/.Uhat/.Vhat/.trup/.state
 * files (fault_output.f90:73-123) */
int wql_iface_set_frames(wql_iface *I, int32_t nslots);
/* copy recorded slots 0..nslots-1 ([slot][39][nz][ny], host) */
int wql_iface_get_frames(wql_iface *I, int32_t nslots, double *out);
/* interface state S[4][nz][ny], DS[4][nz][ny], rupture time [nz][ny] (host) */
int wql_iface_get_state(wql_iface *I, double *S, double *DS, double *trup);
/* rate-and-state state variable psi [nz][ny] (host; I%W of iface_type) */
int wql_iface_get_friction_state(wql_iface *I, double *W);

/* ---- discrete adjoint of the fused stage --------------------------------
 * No reference counterpart (the reference has no adjoint); these serve the
 * gradient of a receiver misfit with respect to per-node rho, lambda, mu.
 * State u = (F, eta), rates r = (DF, Deta); a stage is r = A r + L(u; m),
 * u = u + Bdt r.  wql_adjoint_stage reverses one stage whose INPUT state is
 * the block's current state (restore it with wql_state_load):
 *   rb += Bdt ub;  g += (dL/dm)^T rb;  ub += K^T rb;  rb = A rb
 * with ub, rb the adjoints of u, r, K = dL/du and g the gradient planes.
 * PML slabs, interfaces and the alpha/theta SAT form are not differentiated
 * (wql_adjoint_create fails with any of them); all three precisions are.
 * Dense arrays below are DEVICE pointers, node order x fastest, no ghosts. */
/* adjoint state, rate adjoints, gradients (zeroed) and nslots state slots;
 * WQL_EINVAL for a block with PML slabs, an interface or the alpha/theta
 * SAT form, which the adjoint does not differentiate */
int wql_adjoint_create(wql_block *b, int32_t nslots);
/* the same with nslots_f more slots, numbered nslots .. nslots+nslots_f-1,
 * that hold F only (the memory a slot of F needs, without eta): for the
 * stage inputs of a checkpoint segment, which a reverse stage reads.
 * wql_adjoint_create(b, n) is wql_adjoint_create_f(b, n, 0). */
int wql_adjoint_create_f(wql_block *b, int32_t nslots, int32_t nslots_f);
/* ub = rb = 0, g = 0 */
int wql_adjoint_reset(wql_block *b);
/* copy the current state (F, eta; under FP32-Kahan also F's compensation)
 * into slot / back from slot (checkpoints) */
int wql_state_save(wql_block *b, int32_t slot);
int wql_state_load(wql_block *b, int32_t slot);
/* the same for part of the state: mask WQL_F (F and, under FP32-Kahan, its
 * compensation) and/or WQL_MEMORY (eta; full slots only).  A reverse stage
 * reads only F of its input state, so the stage inputs of a step need only
 * F. */
int wql_state_save_part(wql_block *b, int32_t slot, uint32_t mask);
int wql_state_load_part(wql_block *b, int32_t slot, uint32_t mask);
/* reverse one stage with RK coefficients A (a_s) and Bdt (b_s * dt) */
int wql_adjoint_stage(wql_block *b, double A, double Bdt);
/* ub(F) at the receiver nodes += g[r][0..8] (device array of nrec x 9): the
 * adjoint of one receiver sample */
int wql_adjoint_add_receivers(wql_block *b, const double *g);
/* read / write ub: F part [9][nodes], eta part [6*nmech][nodes] */
int wql_adjoint_get_state(wql_block *b, double *Fbar, double *etabar);
int wql_adjoint_set_state(wql_block *b, const double *Fbar, const double *etabar);
/* accumulated gradients [nodes] each */
int wql_adjoint_gradient(wql_block *b, double *g_rho, double *g_lambda, double *g_mu);

#ifdef __cplusplus
}
#endif
#endif
