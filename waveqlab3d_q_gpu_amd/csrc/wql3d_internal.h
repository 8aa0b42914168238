// Internal types shared by the C-ABI layer (wql3d_api.hip) and the kernels
// (wql3d_kernels.hip).
//
// Device layout of one block (no ghost nodes; node (x,y,z) 0-based, x fastest):
//   idx = x + nx*(y + ny*z);  plane stride P = nx*ny*nz rounded up to a
//   multiple of kPlaneAlign nodes (the tail of each plane is padding that
//   streaming kernels may read and write but no node reads)
//   F[2]  : 9 planes of double each (ping-pong for the fused stage)
//   Fo[2] : FP32-mixed only: 9 float planes each, F[k] rounded to float (the
//           stencil operands of the fused stage)
//   DF    : 9 planes of double
//   G     : geometry planes of type T: cartesian 3 (q_x, r_y, s_z);
//           curvilinear 10 (q_x,r_x,s_x, q_y,r_y,s_y, q_z,r_z,s_z, J)
//   M     : 3 planes (lambda, mu, rho) of type T
//   eta   : 6*nmech planes of T, plane (c6*nmech + i); Deta likewise
//   Qs,Qp : 1 plane each of T
//   pQ/pDQ[f] : 9 planes over the PML slab of face f (slab-local index)
//   face[f]   : 13 planes over the face (l[3], m[3], n[3], |grad xi|, and the
//               metric column of the face axis: xi_x, xi_y, xi_z) on faces
//               with a boundary condition or an interface
// T = double (WQL_FP64) or float (WQL_FP32_MIXED).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/wql3d.h"

namespace wql {

constexpr int kMaxSources = 8;
// support of a mollified (Gaussian, sigma = 2h) source in nodes per axis: at
// 40 nodes the weight is below exp(-50) of its peak even where the spacing
// grows to twice the source's
constexpr int kMollifyRadius = 40;
// z-chunk pipeline of the anelastic fused stage: chunks per stage
constexpr int kMaxChunks = 16;
constexpr int kDefaultChunks = 1;
// plane stride granularity in nodes (see the device layout above)
constexpr long long kPlaneAlign = 256;

// the stencil operands come from the float copy of F (FP32-mixed)
template <typename T, typename S>
constexpr bool kOpCopy = std::is_same<T, float>::value && std::is_same<S, double>::value;
template <typename T, typename S>
using OpType = typename std::conditional<kOpCopy<T, S>, T, S>::type;

struct DevSource {
  int near[3];      // 0-based nearest node
  int order, alpha; // order 2 or 4: singular_source; 0: mollified (Gaussian)
  double off[3];    // location - near_phys, per axis
  double loc[3];    // location
  double am[6];     // A(t) * (mXX, mYY, mZZ, mXY, mXZ, mYZ) for the stage
};

// T: storage type of geometry, material, memory variables and PML fields;
// S: type of the RK state F and rates DF (double, or float with the Kahan
// compensation planes Fc).
template <typename T, typename S = double>
struct StepArgs {
  int nx, ny, nz;
  long long P;      // plane stride (>= nx*ny*nz, see the device layout)
  int nmech;
  double ih[3];     // 1/hq, 1/hr, 1/hs
  double h[3];      // hq, hr, hs
  double pen[3];    // tau0/h per axis (SAT)
  const S* Fin;
  S* Fout;
  S* DF;
  S* Fc;            // Kahan compensation of F (S = float only), updated in place
  // FP32-mixed (T float, S double): F rounded to float, the stencils'
  // operands (each stencil kernel reads F of a node up to 7 times through the
  // caches; the float copy halves those bytes).  Fo_in = (float)F_in exactly,
  // Fo_out is written wherever F_out is.  Null in the other precisions.
  const T* Fo_in;
  T* Fo_out;
  const T* G;
  const double* X;  // node coordinates (3 planes), for mollified sources; else null
  const T* M;
  T* eta;
  T* Deta;
  const T* Qs;
  const T* Qp;
  double tau_inv[8];
  double wgt[8];
  int pml[6];
  int npml[6];
  T* pQ[6];
  T* pDQ[6];
  int bc[6];
  int sat_form;     // WQL_SAT_BC / WQL_SAT_ALPHA_THETA
  const T* face[6];
  // interface SAT forcing of faces with bc 0 coupled to another block:
  // 9 planes over the face, tau0/h * forcing, subtracted from DF like the
  // boundary penalty (null: no interface on the face)
  const double* iface[6];
  int nsrc;
  DevSource src[kMaxSources];
  double A, Bdt;    // fused stage coefficients
  // the rates this stage stores (DF, Deta, PML DQ) are never read: the next
  // stage is the first of a step (A = 0) inside the same wql_time_step call,
  // so the stores are skipped (the call's last step stores them)
  int dead_rates;
  int ix0, ix1, iy0, iy1, iz0, iz1;  // strict interior [lo, hi] 0-based (RHS_Center)
  // node-kernel launch shape (see map_node): y-band height, 256-node segments
  // per band slab, XCD-contiguous remap, non-temporal streams
  int band_rows, segs, xcd_remap, nontemporal;
  // split anelastic stage: the MODE 3 stencil kernels finish the velocity
  // components and write W (12 planes: 6 stress rates, 6 strain rates), then
  // k_memory_update; mech_skip = mechanisms with zero weight whose eta and
  // Deta are identically zero (they stay zero, so they are not streamed)
  T* W;
  int split;
  int mech_skip;
  // strict-interior z-marching kernel (k_interior) of the fused stage: tile
  // shape (0 = off, 1 = 64x8, 2 = 32x16, 3 = 64x4) and z levels per chunk
  int zm;
  int zchunk;
  // box list of k_node: when nbox > 0 the kernel covers these boxes
  // ([x0,x1,y0,y1,z0,z1] inclusive, 0-based) instead of the whole block;
  // box_start[b] = nodes of the boxes before b, box_start[nbox] = total
  int nbox;
  int box[6][6];
  long long box_start[7];
  // flat node range [node0, node1) of k_memory_update (0 and P: every node)
  // and its nodes per thread (1, or 2 with vector accesses when node0 and
  // node1 are even)
  long long node0, node1;
  int mem_npt;
  int mem_nt;       // non-temporal accesses in k_memory_update only
  int mem_xcd;      // workgroup order of k_memory_update_n over the XCDs (xcd_logical)
};

// Launchers (wql3d_kernels.hip).  All return hipError_t of the launch.
enum class RhsPart { kAll = 0, kInterior = 1 };

// fused stage = the strict-interior kernel (if the box is not empty and the
// interior kernel is enabled), the node kernel over the rest of the block
// (the shell), and for anelastic stages with the W hand-off the
// memory-variable kernel over every node
enum class StagePart { kInterior = 0, kShell = 1, kMemory = 2 };
template <typename T, typename S>
hipError_t launch_stage_fused(const StepArgs<T, S>& a, int mesh, hipStream_t s);
template <typename T, typename S>
hipError_t launch_stage_part(const StepArgs<T, S>& a, int mesh, StagePart part, hipStream_t s);
// the same part restricted to the z levels [z0, z1) of the block
template <typename T, typename S>
hipError_t launch_stage_part_z(const StepArgs<T, S>& a, int mesh, StagePart part, int z0, int z1,
                               hipStream_t s);
// true when the stage runs as stencil kernels + k_memory_update (W hand-off)
template <typename T, typename S>
bool stage_has_memory_part(const StepArgs<T, S>& a);
hipError_t launch_rhs(const StepArgs<double>& a, int mesh, RhsPart part, hipStream_t s);
hipError_t launch_sat(const StepArgs<double>& a, hipStream_t s);
hipError_t launch_source(const StepArgs<double>& a, int mesh, hipStream_t s);
template <typename T>
hipError_t launch_scale(T* x, long long n, double A, hipStream_t s);
template <typename T>
hipError_t launch_update(T* x, const T* r, long long n, double inc, hipStream_t s);
template <typename T>
hipError_t launch_fill(T* x, long long n, double v, hipStream_t s);
hipError_t launch_round_float(float* dst, const double* src, long long n, hipStream_t s);
// copy a strided (Fortran-bounded) double array into a dense T plane set and back
template <typename T>
hipError_t launch_gather(T* dst, const double* src, int nx, int ny, int nz, int ncomp,
                         long long sx, long long sy, long long sz, long long sc,
                         long long off0, hipStream_t s);
template <typename T>
hipError_t launch_scatter(double* dst, const T* src, int nx, int ny, int nz, int ncomp,
                          long long sx, long long sy, long long sz, long long sc,
                          long long off0, hipStream_t s);
template <typename S>
hipError_t launch_sample(const S* F, long long P, const long long* idx, int n, double* out,
                         hipStream_t s);
hipError_t launch_metrics(int nx, int ny, int nz, const double* X, double* mx, double* my,
                          double* mz, double* J, hipStream_t s);

// Two-block q-interface: block L's x-high face against block R's x-low face,
// both ny x nz nodes; face index fi = y + ny*z.  One thread per face node
// evaluates both sides (block L's side 2, then block R's side 1), writes the
// SAT forcing of each block, updates the interface state in place and, at
// stage 1 with a slot, the fault output.
constexpr int kFramePlanes = 39;  // Uface 9, Vface 9, S 4, Svel 3, Uhat 6, Vhat 6, trup 1, state 1
template <typename T, typename St = double>
struct IfaceArgs {
  int ny, nz;
  int coupling;
  int nxL, nxR;            // x extents (L's face is x = nxL - 1, R's is x = 0)
  const St* FL;            // block L / R fields at the start of the stage (the
  const St* FR;            // RK state: double, or float under FP32-Kahan)
  long long PL, PR;        // plane strides
  const double* matL;      // lambda, mu, rho of the coupled face nodes (FP64)
  const double* matR;
  const double* frL;       // 13 face planes of L's face 1 and R's face 0 (FP64)
  const double* frR;
  double penL, penR;       // tau0/hq of each block
  double* S;               // 4 planes: slip (3) and slip-path length
  double* DS;              // 4 planes: rates of S
  double* trup;            // rupture time
  // slip-weakening tables: prestress (n, m, l) in the frame of each side,
  // static / dynamic friction coefficient and slip-weakening distance
  const double* T0L;
  const double* T0R;
  const double* fs;
  const double* fd;
  const double* Dc;
  double* satL;            // 9 planes each: tau0/hq * SAT forcing
  double* satR;
  // rate-and-state friction (TPV101 / TPV102 / rate-weakening): state psi
  // and its rate, the parameter a and the nucleation amplitude per node
  // (TPV101: 25 F(r); rate-weakening: 2 F(r), times its -0.2429 T0_n), and
  // the constants L0, f0, b, V0
  double* W;
  double* DW;
  const double* rs_a;
  const double* rs_nuc;
  // rate-weakening: the flash-weakening velocity V_w per node (else null)
  const double* rs_vw;
  double rs_L0, rs_f0, rs_b, rs_V0;
  // linear friction: tau = alpha v
  double alpha;
  double A, Bdt, t;
  int stage1;
  double* frame;           // null, or kFramePlanes planes of the current output slot
};
template <typename T, typename S>
hipError_t launch_iface(const IfaceArgs<T, S>& a, hipStream_t s);

}  // namespace wql
