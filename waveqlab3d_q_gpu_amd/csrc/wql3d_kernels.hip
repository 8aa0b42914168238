// HIP kernels for gfx950 (MI355X) of the per-block WaveQLab3D-Q time step.
//
// A fused RK stage is three kernels: k_interior (the strict interior, a
// z-marching 64x8 tile with its x/y operands staged in LDS and a register
// queue along z), k_node (the shell: closure rows, PML slabs, faces; one
// thread per node) and k_memory_update_n (memory variables and the stress
// components' RK update, pure streaming, two nodes per thread).
//
// k_node: the workgroup (256 threads = 4 wavefronts) covers a contiguous run
// of one box row by row; x is the fastest index, so every plane load of a
// wavefront is unit stride.  Stencil neighbours are re-read through the L1 /
// L2 caches; the per-node streams (DF, memory variables, PML auxiliaries)
// are read and written exactly once per stage.  Closure rows differ only per
// (y,z) line for the y/z axes, so wavefronts stay convergent there.  Device
// helpers: wql3d_device.h.
#include "wql3d_device.h"

namespace wql {

// ---------------------------------------------------------------------------
// The per-node kernel.
//   MODE 0: fused RK stage   DF = A*DF + rates;  F_out = F_in + Bdt*DF
//   MODE 1: rates only       DF += rates  (set_rates_elastic, all nodes)
//   MODE 2: rates only, strict interior nodes (RHS_Center)
// In MODE 0 rates include the moment-tensor source and the SAT penalty.
// ---------------------------------------------------------------------------
// Logical order is band-major: bands of band_rows y-rows, then z, then
// 256-node segments of the band's (x,y) slab.
template <typename T, typename S>
__device__ __forceinline__ bool map_node(const StepArgs<T, S>& a, int& x, int& y, int& z) {
  const unsigned L = xcd_logical(blockIdx.x, gridDim.x, a.xcd_remap);
  const unsigned seg = L % (unsigned)a.segs;
  const unsigned bz = L / (unsigned)a.segs;
  z = (int)(bz % (unsigned)a.nz);
  const int band = (int)(bz / (unsigned)a.nz);
  const int tb = (int)seg * 256 + (int)threadIdx.x;
  const int yb = tb / a.nx;
  x = tb - yb * a.nx;
  y = band * a.band_rows + yb;
  return yb < a.band_rows && y < a.ny;
}

template <typename T, int MESH, int NMECH, int MODE, bool NT, bool SHELL, typename S>
__device__ __forceinline__ void node_epilogue(const StepArgs<T, S>& a, int x, int y, int z,
                                              long long id, const T (&U)[3][9]);

// Node of the shell boxes (nbox > 0): the global thread index runs over the
// concatenated boxes, x fastest inside each box.
template <typename T, typename S>
__device__ __forceinline__ bool map_box_node(const StepArgs<T, S>& a, int& x, int& y, int& z) {
  const unsigned h = xcd_logical(blockIdx.x, gridDim.x, a.xcd_remap);
  const long long g = (long long)h * 256 + threadIdx.x;
  if (g >= a.box_start[a.nbox]) return false;
  int b = 0;
  while (b + 1 < a.nbox && g >= a.box_start[b + 1]) ++b;
  const long long t = g - a.box_start[b];
  const int bx = a.box[b][1] - a.box[b][0] + 1, by = a.box[b][3] - a.box[b][2] + 1;
  const long long row = t / bx;
  x = a.box[b][0] + (int)(t - row * bx);
  y = a.box[b][2] + (int)(row % by);
  z = a.box[b][4] + (int)(row / by);
  return true;
}

template <typename T, int MESH, int NMECH, int MODE, bool NT, typename S>
__global__ __launch_bounds__(256) void k_node(const StepArgs<T, S> a) {
  int x, y, z;
  if (a.nbox > 0) {
    if (!map_box_node(a, x, y, z)) return;
  } else if (!map_node(a, x, y, z)) {
    return;
  }
  if (MODE == 2) {
    if (x < a.ix0 || x > a.ix1 || y < a.iy0 || y > a.iy1 || z < a.iz0 || z > a.iz1) return;
  }
  const long long plane = (long long)a.nx * a.ny;
  const long long id = x + (long long)a.nx * y + (long long)z * plane;
  const int pos[3] = {x, y, z};

  T U[3][9];
  if (MESH) derivs_curv<T>(a, pos, id, U);
  else derivs_cart<T>(a, pos, id, U);
  node_epilogue<T, MESH, NMECH, MODE, NT, true>(a, x, y, z, id, U);
}

// Everything after the derivatives at one node: elastic rates, memory
// variables (MODE 0/1/2) or the W hand-off (MODE 3), PML, SAT, source and
// the RK update or rate accumulation.  SHELL = false is the strict interior
// (RHS_Center): no PML slab and no face, so those paths are compiled out.
template <typename T, int MESH, int NMECH, int MODE, bool NT, bool SHELL, typename S>
__device__ __forceinline__ void node_epilogue(const StepArgs<T, S>& a, int x, int y, int z,
                                              long long id, const T (&U)[3][9]) {
  const long long P = a.P;
  const unsigned idu = (unsigned)id;
  const T lam = ld_stream<NT>(atp(a.M, idu)), mu = ld_stream<NT>(atp(a.M + P, idu)),
          rho = ld_stream<NT>(atp(a.M + 2 * P, idu));
  const T Jp = MESH ? a.G[9 * P + id] : T(1);
  const T rhoJ_inv = T(1) / (rho * Jp);
  const T l2m = lam + T(2) * mu;

  // elastic rates (RHS_Interior.f90:238-254)
  T DU[9];
  DU[0] = (U[0][3] + U[1][6] + U[2][7]) * rhoJ_inv;
  DU[1] = (U[0][6] + U[1][4] + U[2][8]) * rhoJ_inv;
  DU[2] = (U[0][7] + U[1][8] + U[2][5]) * rhoJ_inv;
  DU[3] = l2m * U[0][0] + lam * (U[1][1] + U[2][2]);
  DU[4] = l2m * U[1][1] + lam * (U[0][0] + U[2][2]);
  DU[5] = l2m * U[2][2] + lam * (U[0][0] + U[1][1]);
  DU[6] = mu * (U[1][0] + U[0][1]);
  DU[7] = mu * (U[2][0] + U[0][2]);
  DU[8] = mu * (U[2][1] + U[1][2]);

  // PML slab membership, one face per axis (pml_damping_and_q)
  int fa[3] = {-1, -1, -1};
  long long sidx[3] = {0, 0, 0};
  T dd[3] = {T(0), T(0), T(0)};
  if (MODE != 2 && SHELL) {
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      if (!a.pml[f]) continue;
      long long si;
      double d;
      if (pml_slab(a, f, x, y, z, &si, &d)) {
        fa[f >> 1] = f;
        sidx[f >> 1] = si;
        dd[f >> 1] = (T)d;
      }
    }
  }
  const bool inpml = fa[0] >= 0 || fa[1] >= 0 || fa[2] >= 0;

  // anelastic Q4/Q8 (RHS_Interior.f90:2599-2906; interior JU_xJU_yJU_z6.f90:16299-16376)
  constexpr bool FUSED = (MODE == 0 || MODE == 3);
  T e[6];
  if (NMECH > 0) {
    e[0] = U[0][0];
    e[1] = U[1][1];
    e[2] = U[2][2];
    e[3] = U[1][0] + U[0][1];
    e[4] = U[2][0] + U[0][2];
    e[5] = U[2][1] + U[1][2];
    if (inpml) {
      T qx[3] = {T(0), T(0), T(0)}, qy[3] = {T(0), T(0), T(0)}, qz[3] = {T(0), T(0), T(0)};
      if (fa[0] >= 0) {
        const T* Q = a.pQ[fa[0]];
        const long long Ps = slab_plane(a, fa[0]);
        qx[0] = Q[3 * Ps + sidx[0]]; qx[1] = Q[6 * Ps + sidx[0]]; qx[2] = Q[7 * Ps + sidx[0]];
      }
      if (fa[1] >= 0) {
        const T* Q = a.pQ[fa[1]];
        const long long Ps = slab_plane(a, fa[1]);
        qy[0] = Q[4 * Ps + sidx[1]]; qy[1] = Q[6 * Ps + sidx[1]]; qy[2] = Q[8 * Ps + sidx[1]];
      }
      if (fa[2] >= 0) {
        const T* Q = a.pQ[fa[2]];
        const long long Ps = slab_plane(a, fa[2]);
        qz[0] = Q[5 * Ps + sidx[2]]; qz[1] = Q[7 * Ps + sidx[2]]; qz[2] = Q[8 * Ps + sidx[2]];
      }
      e[0] = e[0] - dd[0] * qx[0];
      e[1] = e[1] - dd[1] * qy[0];
      e[2] = e[2] - dd[2] * qz[0];
      e[3] = e[3] - dd[0] * qx[1] - dd[1] * qy[1];
      e[4] = e[4] - dd[0] * qx[2] - dd[2] * qz[1];
      e[5] = e[5] - dd[1] * qy[2] - dd[2] * qz[2];
    }
  }
  if (NMECH > 0 && MODE != 3) {
    const T tr = e[0] + e[1] + e[2];
    const T Qs = ld_stream<NT>(atp(a.Qs, idu)), Qp = ld_stream<NT>(atp(a.Qp, idu));
    const T twomuQs = T(2) * mu * Qs;
    const T pfac = l2m * Qp - twomuQs;
    const T muQs = mu * Qs;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      T ev[NMECH > 0 ? NMECH : 1];
      T sum = T(0);
#pragma unroll
      for (int i = 0; i < NMECH; ++i) {
        ev[i] = ld_stream<NT>(atp(a.eta + (long long)(c * NMECH + i) * P, idu));
        sum += ev[i];
      }
      DU[3 + c] -= sum;
#pragma unroll
      for (int i = 0; i < NMECH; ++i) {
        const T w = (T)a.wgt[i];
        const T ti = (T)a.tau_inv[i];
        const T src = c < 3 ? (w * twomuQs) * e[c] + (w * pfac) * tr : (w * muQs) * e[c];
        const T r = (src - ev[i]) * ti;
        T* De = atp(a.Deta + (long long)(c * NMECH + i) * P, idu);
        T* Et = atp(a.eta + (long long)(c * NMECH + i) * P, idu);
        if (MODE == 0) {
          const T dn = scale_rate(a.A, ld_rate<NT>(a.A, De)) + r;
          if (!a.dead_rates) st_stream<NT>(De, dn);
          st_stream<NT>(Et, ev[i] + (T)a.Bdt * dn);
        } else {
          *De = *De + r;
        }
      }
    }
  }

  // SAT penalty terms at face nodes (impose_boundary_condition) and the
  // interface forcing of coupled faces (Impose_Interface_Condition,
  // RHS_Interior.f90:1813-1836, computed by k_iface), fused mode
  double satDF[9];
  bool has_sat = false;
  if (FUSED && SHELL) {
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      if (a.bc[f] < 0 || (a.bc[f] == 0 && !a.iface[f])) continue;
      long long fidx, NF;
      if (!on_face(a, f, x, y, z, &fidx, &NF)) continue;
      if (!has_sat) {
#pragma unroll
        for (int c = 0; c < 9; ++c) satDF[c] = 0.0;
        has_sat = true;
      }
      if (a.bc[f] == 0) {
#pragma unroll
        for (int c = 0; c < 9; ++c) satDF[c] += a.iface[f][c * NF + fidx];
        continue;
      }
      T u[9], ubc[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) u[c] = (T)a.Fin[c * P + id];
      const T pen = (T)a.pen[f >> 1];
      T Ux[9];
      sat_rate(a, f, fidx, NF, u, lam, mu, rho, ubc, Ux);
#pragma unroll
      for (int c = 0; c < 9; ++c) satDF[c] += (double)(pen * Ux[c]);
    }
  }

  // PML feedback and auxiliary rates (RHS_Interior.f90:1846-2265)
  if (inpml) {
    const T cfs = T(1);
#pragma unroll
    for (int ax = 2; ax >= 0; --ax) {   // reference order: z, y, x slabs
      const int f = fa[ax];
      if (f < 0) continue;
      const T d = dd[ax];
      T* Q = a.pQ[f];
      T* DQ = a.pDQ[f];
      const long long Ps = slab_plane(a, f);
      const long long si = sidx[ax];
      T q[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) q[c] = ld_stream<NT>(Q + c * Ps + si);
      DU[0] -= d * Jp * q[0] * rhoJ_inv;
      DU[1] -= d * Jp * q[1] * rhoJ_inv;
      DU[2] -= d * Jp * q[2] * rhoJ_inv;
      DU[3] -= d * (l2m * q[3] + lam * q[4] + lam * q[5]);
      DU[4] -= d * (lam * q[3] + l2m * q[4] + lam * q[5]);
      DU[5] -= d * (lam * q[3] + lam * q[4] + l2m * q[5]);
      DU[6] -= mu * d * q[6];
      DU[7] -= mu * d * q[7];
      DU[8] -= mu * d * q[8];
      T s[9];
      const T iJ = T(1) / Jp;
      if (ax == 0) {
        s[0] = U[0][3] * iJ; s[1] = U[0][6] * iJ; s[2] = U[0][7] * iJ;
        s[3] = U[0][0]; s[4] = T(0); s[5] = T(0); s[6] = U[0][1]; s[7] = U[0][2]; s[8] = T(0);
      } else if (ax == 1) {
        s[0] = U[1][6] * iJ; s[1] = U[1][4] * iJ; s[2] = U[1][8] * iJ;
        s[3] = T(0); s[4] = U[1][1]; s[5] = T(0); s[6] = U[1][0]; s[7] = T(0); s[8] = U[1][2];
      } else {
        s[0] = U[2][7] * iJ; s[1] = U[2][8] * iJ; s[2] = U[2][5] * iJ;
        s[3] = T(0); s[4] = T(0); s[5] = U[2][2]; s[6] = T(0); s[7] = U[2][0]; s[8] = U[2][1];
      }
      // SAT on the slab's own face adds to this slab's DQ (RHS_Interior.f90:1640-1643)
      T sat[9];
      bool fsat = false;
      if (FUSED && a.bc[f] > 0) {
        long long fidx, NF;
        if (on_face(a, f, x, y, z, &fidx, &NF)) {
          T u[9];
#pragma unroll
          for (int c = 0; c < 9; ++c) u[c] = (T)a.Fin[c * P + id];
          sat_ubc(a, f, fidx, NF, u, lam, mu, rho, sat);
          fsat = true;
        }
      }
      const T pen = (T)a.pen[ax];
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        T r = s[c] - (d + cfs) * q[c];
        if (fsat) r -= pen * sat[c];
        if (FUSED) {
          const T dn = scale_rate(a.A, ld_rate<NT>(a.A, DQ + c * Ps + si)) + r;
          if (!a.dead_rates) st_stream<NT>(DQ + c * Ps + si, dn);
          st_stream<NT>(Q + c * Ps + si, q[c] + (T)a.Bdt * dn);
        } else {
          DQ[c * Ps + si] = DQ[c * Ps + si] + r;
        }
      }
    }
  }

  if (MODE == 0) {
    double srcv[6];
    const bool has_src = a.nsrc > 0 && source_at<T, MESH>(a, x, y, z, id, srcv);
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      double df = scale_rate(a.A, ld_rate<NT>(a.A, atp(a.DF + c * P, idu))) + (double)DU[c];
      if (has_src && c >= 3) df -= srcv[c - 3];
      if (has_sat) df -= satDF[c];
      rk_finish<NT>(a, c, idu, df);
    }
  } else if (MODE == 3) {
    // the velocity rates carry no memory-variable term: finish their RK
    // update here; hand the stress rates without memory-variable terms and
    // the (PML-modified) strain rates to k_memory_update through W
    double srcv[6];
    const bool has_src = a.nsrc > 0 && source_at<T, MESH>(a, x, y, z, id, srcv);
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      T r = DU[c];
      if (has_src && c >= 3) r -= (T)srcv[c - 3];
      if (has_sat) r -= (T)satDF[c];
      if (c < 3) {
        rk_finish<NT>(a, c, idu, scale_rate(a.A, ld_rate<NT>(a.A, atp(a.DF + c * P, idu))) + (double)r);
      } else {
        st_stream<NT>(atp(a.W + (c - 3) * P, idu), r);
      }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) st_stream<NT>(atp(a.W + (6 + c) * P, idu), e[c]);
  } else {
#pragma unroll
    for (int c = 0; c < 9; ++c) a.DF[c * P + id] = a.DF[c * P + id] + (double)DU[c];
  }
}

// Memory-variable part of a fused anelastic stage, one thread per node over
// the flat node index (pure streaming, unit stride): W holds the stress rates
// without the memory-variable terms (planes 0..5) and the strain rates used
// by the relaxation equations (planes 6..11), both from the MODE 3 stencil
// kernels, which also finish the velocity components.
//   DU(4..9) -= sum_i eta_i;  Deta_i = A*Deta_i + (src_i(eps) - eta_i)/tau_i;
//   eta_i += Bdt*Deta_i;  DF(4..9) = A*DF + DU;  F_out(4..9) = F_in + Bdt*DF
// Memory variables of one node in a fused stage: DU(4..9) -= sum_i eta_i and
// the relaxation update of eta_i / Deta_i from the strain rates e
// (RHS_Interior.f90:2599-2906).  Mechanisms in mech_skip are not streamed.
// The memory-variable update proper.  eta and Deta are passed as restrict
// pointers so that the loads of one component can be issued ahead of the
// stores of the previous one; each component loads all of its mechanisms
// before it stores any.
template <typename T, int NMECH, bool NT>
__device__ __forceinline__ void memory_update_planes(T* __restrict__ eta, T* __restrict__ Deta,
                                                     const T* __restrict__ Qsp,
                                                     const T* __restrict__ Qpp, long long P,
                                                     int mech_skip, const double* wgt,
                                                     const double* tau_inv, double A, double Bdt,
                                                     unsigned id, T mu, T l2m, const T (&e)[6],
                                                     T (&DU)[9], bool store_rates = true) {
  const T tr = e[0] + e[1] + e[2];
  const T Qs = ld_stream<NT>(atp(Qsp, id)), Qp = ld_stream<NT>(atp(Qpp, id));
  const T twomuQs = T(2) * mu * Qs;
  const T pfac = l2m * Qp - twomuQs;
  const T muQs = mu * Qs;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T ev[NMECH], dv[NMECH];
    T sum = T(0);
#pragma unroll
    for (int i = 0; i < NMECH; ++i) {
      if (mech_skip & (1 << i)) { ev[i] = T(0); dv[i] = T(0); continue; }
      ev[i] = ld_stream<NT>(atp(eta + (long long)(c * NMECH + i) * P, id));
      dv[i] = ld_rate<NT>(A, atp(Deta + (long long)(c * NMECH + i) * P, id));
      sum += ev[i];
    }
    DU[3 + c] -= sum;
#pragma unroll
    for (int i = 0; i < NMECH; ++i) {
      if (mech_skip & (1 << i)) continue;
      const T w = (T)wgt[i];
      const T ti = (T)tau_inv[i];
      const T src = c < 3 ? (w * twomuQs) * e[c] + (w * pfac) * tr : (w * muQs) * e[c];
      const T r = (src - ev[i]) * ti;
      const T dn = scale_rate(A, dv[i]) + r;
      if (store_rates) st_stream<NT>(atp(Deta + (long long)(c * NMECH + i) * P, id), dn);
      st_stream<NT>(atp(eta + (long long)(c * NMECH + i) * P, id), ev[i] + (T)Bdt * dn);
    }
  }
}

template <typename T, int NMECH, bool NT, typename S>
__device__ __forceinline__ void memory_update_node(const StepArgs<T, S>& a, unsigned id, T mu, T l2m,
                                                   const T (&e)[6], T (&DU)[9]) {
  memory_update_planes<T, NMECH, NT>(a.eta, a.Deta, a.Qs, a.Qp, a.P, a.mech_skip, a.wgt, a.tau_inv,
                                     a.A, a.Bdt, id, mu, l2m, e, DU, !a.dead_rates);
}

// Nodes of the strict interior outside every source box: no PML, face or
// source term enters their stress rates, which are therefore Hooke's law of
// the strain rates in W.  The interior kernel stores only the strain rates
// for them, and the memory-variable kernels form the stress rates from those
// (the same expressions as the stencil epilogue).  The test is geometric, so
// it holds whichever kernel produced W.
template <typename T, typename S>
__device__ __forceinline__ bool hooke_node(const StepArgs<T, S>& a, unsigned long long id) {
  const unsigned long long plane = (unsigned long long)a.nx * (unsigned long long)a.ny;
  const int z = (int)(id / plane);
  const unsigned long long r = id - (unsigned long long)z * plane;
  const int y = (int)(r / (unsigned long long)a.nx);
  const int x = (int)(r - (unsigned long long)y * a.nx);
  if (x < a.ix0 || x > a.ix1 || y < a.iy0 || y > a.iy1 || z < a.iz0 || z > a.iz1) return false;
  for (int k = 0; k < a.nsrc; ++k) {
    const DevSource& Sr = a.src[k];
    const int rr = source_radius(Sr);   // the box of source_at
    if (x >= Sr.near[0] - rr && x <= Sr.near[0] + rr && y >= Sr.near[1] - rr &&
        y <= Sr.near[1] + rr && z >= Sr.near[2] - rr && z <= Sr.near[2] + rr)
      return false;
  }
  return true;
}

template <typename T>
__device__ __forceinline__ void hooke_rates(T lam, T l2m, T mu, const T (&e)[6], T* DU) {
  DU[3] = l2m * e[0] + lam * (e[1] + e[2]);
  DU[4] = l2m * e[1] + lam * (e[0] + e[2]);
  DU[5] = l2m * e[2] + lam * (e[0] + e[1]);
  DU[6] = mu * e[3];
  DU[7] = mu * e[4];
  DU[8] = mu * e[5];
}

template <typename T, int NMECH, bool NT, typename S>
__global__ __launch_bounds__(256) void k_memory_update(const StepArgs<T, S> a) {
  const unsigned id = (unsigned)a.node0 + blockIdx.x * 256u + threadIdx.x;
  const long long P = a.P;
  if (id >= (unsigned long long)a.node1) return;
  T DU[9], e[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) DU[c] = T(0);
#pragma unroll
  for (int c = 0; c < 6; ++c) e[c] = ld_stream<NT>(atp(a.W + (6 + c) * P, id));
  const T lam = ld_stream<NT>(atp(a.M, id)), mu = ld_stream<NT>(atp(a.M + P, id));
  const T l2m = lam + T(2) * mu;
  if (hooke_node(a, id)) {
    hooke_rates(lam, l2m, mu, e, DU);
  } else {
#pragma unroll
    for (int c = 3; c < 9; ++c) DU[c] = ld_stream<NT>(atp(a.W + (c - 3) * P, id));
  }
  memory_update_node<T, NMECH, NT>(a, id, mu, l2m, e, DU);
#pragma unroll
  for (int c = 3; c < 9; ++c) {
    const double df = scale_rate(a.A, ld_rate<NT>(a.A, atp(a.DF + c * P, id))) + (double)DU[c];
    rk_finish<NT>(a, c, id, df);
  }
}

// The same update with N consecutive nodes per thread (N = 2): every plane
// access is one N-wide vector load or store, so a wavefront moves twice the
// bytes per memory instruction; the arithmetic per node is unchanged (same
// operations in the same order as memory_update_planes / rk_finish).
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
typedef double d4v __attribute__((ext_vector_type(4)));
template <int N, typename V> struct VecN;
template <> struct VecN<2, float> { using type = f2v; };
template <> struct VecN<2, double> { using type = d2v; };
template <> struct VecN<4, float> { using type = f4v; };
template <> struct VecN<4, double> { using type = d4v; };

// N-wide plane access at node id (id a multiple of N); NT = non-temporal
template <int N, bool NT = false, typename V>
__device__ __forceinline__ void ldn(const V* base, unsigned id, V (&o)[N]) {
  using VV = typename VecN<N, V>::type;
  const VV* p = reinterpret_cast<const VV*>(reinterpret_cast<const char*>(base) +
                                            id * (unsigned)sizeof(V));
  VV t;
  if constexpr (NT) t = __builtin_nontemporal_load(p);
  else t = *p;
#pragma unroll
  for (int k = 0; k < N; ++k) o[k] = t[k];
}
template <int N, bool NT = false, typename V>
__device__ __forceinline__ void stn(V* base, unsigned id, const V (&v)[N]) {
  using VV = typename VecN<N, V>::type;
  VV* p = reinterpret_cast<VV*>(reinterpret_cast<char*>(base) + id * (unsigned)sizeof(V));
  VV t;
#pragma unroll
  for (int k = 0; k < N; ++k) t[k] = v[k];
  if constexpr (NT) __builtin_nontemporal_store(t, p);
  else *p = t;
}

template <typename T, int NMECH, int N, bool NT, bool RD, bool WR>
__device__ __forceinline__ void memory_update_planes_n(
    T* __restrict__ eta, T* __restrict__ Deta, const T* __restrict__ Qsp, const T* __restrict__ Qpp,
    long long P, int mech_skip, const double* wgt, const double* tau_inv, double A, double Bdt,
    unsigned id, const T (&mu)[N], const T (&l2m)[N], const T (&e)[N][6], T (&DU)[N][9]) {
  T Qs[N], Qp[N];
  ldn<N, NT>(Qsp, id, Qs);
  ldn<N, NT>(Qpp, id, Qp);
  T tr[N], twomuQs[N], pfac[N], muQs[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    tr[n] = e[n][0] + e[n][1] + e[n][2];
    twomuQs[n] = T(2) * mu[n] * Qs[n];
    pfac[n] = l2m[n] * Qp[n] - twomuQs[n];
    muQs[n] = mu[n] * Qs[n];
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T ev[NMECH][N], dv[NMECH][N];
    T sum[N];
#pragma unroll
    for (int n = 0; n < N; ++n) sum[n] = T(0);
#pragma unroll
    for (int i = 0; i < NMECH; ++i) {
      if (mech_skip & (1 << i)) {
#pragma unroll
        for (int n = 0; n < N; ++n) { ev[i][n] = T(0); dv[i][n] = T(0); }
        continue;
      }
      ldn<N, NT>(eta + (long long)(c * NMECH + i) * P, id, ev[i]);
      if constexpr (RD) ldn<N, NT>(Deta + (long long)(c * NMECH + i) * P, id, dv[i]);
      else
#pragma unroll
        for (int n = 0; n < N; ++n) dv[i][n] = T(0);
#pragma unroll
      for (int n = 0; n < N; ++n) sum[n] += ev[i][n];
    }
#pragma unroll
    for (int n = 0; n < N; ++n) DU[n][3 + c] -= sum[n];
#pragma unroll
    for (int i = 0; i < NMECH; ++i) {
      if (mech_skip & (1 << i)) continue;
      const T w = (T)wgt[i];
      const T ti = (T)tau_inv[i];
      T dn[N], en[N];
#pragma unroll
      for (int n = 0; n < N; ++n) {
        const T src = c < 3 ? (w * twomuQs[n]) * e[n][c] + (w * pfac[n]) * tr[n]
                            : (w * muQs[n]) * e[n][c];
        const T r = (src - ev[i][n]) * ti;
        dn[n] = scale_rate(A, dv[i][n]) + r;
        en[n] = ev[i][n] + (T)Bdt * dn[n];
      }
      if constexpr (WR) stn<N, NT>(Deta + (long long)(c * NMECH + i) * P, id, dn);
      stn<N, NT>(eta + (long long)(c * NMECH + i) * P, id, en);
    }
  }
}

// RD = false at the first RK stage (A = 0): the old rates Deta and DF are
// not read (scale_rate maps them to 0 there whatever they hold).  WR = false
// when the stage's rates are dead (StepArgs::dead_rates): Deta and DF are
// not stored.
template <typename T, int NMECH, int N, bool NT, typename S, bool RD, bool WR>
__global__ __launch_bounds__(256) void k_memory_update_n(const StepArgs<T, S> a) {
  const unsigned wg = xcd_logical(blockIdx.x, gridDim.x, a.mem_xcd);
  const unsigned id = (unsigned)a.node0 + (wg * 256u + threadIdx.x) * (unsigned)N;
  const long long P = a.P;
  if (id >= (unsigned long long)a.node1) return;
  T DU[N][9], e[N][6];
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int c = 0; c < 3; ++c) DU[n][c] = T(0);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    T t[N];
    ldn<N, NT>(a.W + (6 + c) * P, id, t);
#pragma unroll
    for (int n = 0; n < N; ++n) e[n][c] = t[n];
  }
  T lam[N], mu[N], l2m[N];
  ldn<N, NT>(a.M, id, lam);
  ldn<N, NT>(a.M + P, id, mu);
#pragma unroll
  for (int n = 0; n < N; ++n) l2m[n] = lam[n] + T(2) * mu[n];
  bool hk[N], all_hooke = true;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    hk[n] = hooke_node(a, (unsigned long long)id + n);
    all_hooke = all_hooke && hk[n];
  }
  if (!all_hooke) {
#pragma unroll
    for (int c = 3; c < 9; ++c) {
      T t[N];
      ldn<N, NT>(a.W + (c - 3) * P, id, t);
#pragma unroll
      for (int n = 0; n < N; ++n) DU[n][c] = t[n];
    }
  }
#pragma unroll
  for (int n = 0; n < N; ++n)
    if (hk[n]) hooke_rates(lam[n], l2m[n], mu[n], e[n], DU[n]);
  memory_update_planes_n<T, NMECH, N, NT, RD, WR>(a.eta, a.Deta, a.Qs, a.Qp, P, a.mech_skip, a.wgt,
                                      a.tau_inv, a.A, a.Bdt, id, mu, l2m, e, DU);
#pragma unroll
  for (int c = 3; c < 9; ++c) {
    S dfo[N], fin[N];
    if constexpr (RD) ldn<N, NT>(a.DF + c * P, id, dfo);
    else
#pragma unroll
      for (int n = 0; n < N; ++n) dfo[n] = S(0);
    ldn<N, NT>(a.Fin + c * P, id, fin);
    if constexpr (std::is_same<S, double>::value) {
      double df[N], fo[N];
#pragma unroll
      for (int n = 0; n < N; ++n) {
        df[n] = scale_rate(a.A, dfo[n]) + (double)DU[n][c];
        fo[n] = fin[n] + a.Bdt * df[n];
      }
      if constexpr (WR) stn<N, NT>(a.DF + c * P, id, df);
      stn<N, NT>(a.Fout + c * P, id, fo);
      if constexpr (kOpCopy<T, S>) {
        T fo32[N];
#pragma unroll
        for (int n = 0; n < N; ++n) fo32[n] = (T)fo[n];
        stn<N, NT>(a.Fo_out + c * P, id, fo32);
      }
    } else {
      float d[N], cn[N], fo[N], cpv[N];
      ldn<N, NT>(a.Fc + c * P, id, cpv);
#pragma unroll
      for (int n = 0; n < N; ++n) {
        d[n] = (float)(scale_rate(a.A, dfo[n]) + (double)DU[n][c]);
        const float y = (float)a.Bdt * d[n] - cpv[n];
        const float t = fin[n] + y;
        cn[n] = (t - fin[n]) - y;
        fo[n] = t;
      }
      if constexpr (WR) stn<N, NT>(a.DF + c * P, id, d);
      stn<N, NT>(a.Fc + c * P, id, cn);
      stn<N, NT>(a.Fout + c * P, id, fo);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused stage over the strict interior box (the RHS_Center region,
// JU_xJU_yJU_z6.f90:16158-16173: no closure row, no PML slab, no face), as a
// z-marching kernel in flux form.  Per node and direction d the stencil
// operands are the velocities v_c and the fluxes
//   curvilinear  fl_d[c] = J * sum_m X_m(d) * sigma(m,c)
//   Cartesian    fl_d[c] = sigma(d,c)
// so that
//   rho*J * dv_c/dt = sum_d 1/h_d D_d(fl_d[c])          (Cartesian: X_d/h_d)
//   dv_c/dx_m       = sum_d X_m(pt,d)/h_d D_d(v_c)
// which are the elastic rates of RHS_Interior.f90:238-254 with the three
// metric terms of JJU_x6 summed before, instead of after, differencing.
//
// A workgroup owns a TX x TY tile of the (x,y) plane and a chunk of z levels.
// Per level, v and fl_q / fl_r of the tile plus a 3-node halo are staged in
// LDS for the x and y derivatives; the z derivative reads a 7-level register
// queue of (v, fl_s) of the thread's own column.  The epilogue applies the
// memory variables, the moment-tensor source and the RK update in place, so
// each per-node stream is read and written once.
// ---------------------------------------------------------------------------
// stress component index in F of sigma(m,c): F(3..8) = sxx, syy, szz, sxy, sxz, syz
__host__ __device__ constexpr int sig_comp(int m, int c) {
  return m == c ? 3 + m : (m + c == 1 ? 6 : (m + c == 2 ? 7 : 8));
}

// operands of node p for direction D: v (3) and fl_D (3).  load_ops issues
// the loads (F in fp64, J and X_m(D) in the storage type), make_ops forms the
// operands, so the two can be placed apart to hide the load latency.
template <typename T, int MESH, typename S = double>
struct RawOps {
  OpType<T, S> f[9];
  T g[MESH ? 4 : 1];
};

template <typename T, int MESH, int D, typename S>
__device__ __forceinline__ void load_ops(const StepArgs<T, S>& a, unsigned p,
                                         RawOps<T, MESH, S>& r) {
  const long long P = a.P;
#pragma unroll
  for (int c = 0; c < 9; ++c) r.f[c] = ldp(op_planes(a) + c * P, p);
  if (MESH) {
    r.g[0] = ldp(a.G + 9 * P, p);
#pragma unroll
    for (int m = 0; m < 3; ++m) r.g[1 + m] = ldp(a.G + (3 * m + D) * P, p);
  }
}

template <typename T, int MESH, int D, typename S>
__device__ __forceinline__ void make_ops(const RawOps<T, MESH, S>& r, T (&v)[3], T (&fl)[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = (T)r.f[c];
  if (MESH) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
      fl[c] = r.g[0] * (r.g[1] * (T)r.f[sig_comp(0, c)] + r.g[2] * (T)r.f[sig_comp(1, c)] +
                        r.g[3] * (T)r.f[sig_comp(2, c)]);
  } else {
#pragma unroll
    for (int c = 0; c < 3; ++c) fl[c] = (T)r.f[sig_comp(D, c)];
  }
}

template <typename T, int MESH, int D, typename S>
__device__ __forceinline__ void node_operands(const StepArgs<T, S>& a, unsigned p, T (&v)[3],
                                              T (&fl)[3]) {
  RawOps<T, MESH, S> r;
  load_ops<T, MESH, D>(a, p, r);
  make_ops<T, MESH, D>(r, v, fl);
}

// An empty asm statement that takes v in a VGPR and hands it back: the value
// must be computed at this point of the program.
template <typename V>
__device__ __forceinline__ void pin(V& v) {
  asm volatile("" : "+v"(v));
}

// minimum waves per SIMD the interior kernel is compiled for (register cap
// 512 / waves): fp32 storage / fp64 storage
#ifndef WQL_IW32
#define WQL_IW32 2
#endif
#ifndef WQL_IW64
#define WQL_IW64 2
#endif
#define WQL_INTERIOR_WAVES(T) (sizeof(T) == 4 ? WQL_IW32 : WQL_IW64)
// epilogue operands of the FP32 interior kernel loaded ahead (1) or in the
// epilogue (0)
#ifndef WQL_IPRE
#define WQL_IPRE 1
#endif

// MODE 0: fused stage (memory variables in the epilogue); MODE 3: rates and
// strain rates to W, memory variables and RK update by k_memory_update
template <typename T, int MESH, int NMECH, int MODE, bool NT, int TX, int TY, typename S>
__global__ __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(WQL_INTERIOR_WAVES(T))))
void k_interior(const StepArgs<T, S> a) {
  constexpr int HX = TX + 6, HY = TY + 6, NTH = TX * TY;
  // planes: v (0..2), fl_q (3..5), fl_r (6..8)
  __shared__ T L[9][HY][HX];

  const int bx = a.ix1 - a.ix0 + 1, by = a.iy1 - a.iy0 + 1;
  const unsigned ntx = (unsigned)((bx + TX - 1) / TX), nty = (unsigned)((by + TY - 1) / TY);
  const unsigned Lidx = xcd_logical(blockIdx.x, gridDim.x, a.xcd_remap);
  const int tix = (int)(Lidx % ntx);
  const int tiy = (int)((Lidx / ntx) % nty);
  const int zc = (int)(Lidx / (ntx * nty));
  const int z0 = a.iz0 + zc * a.zchunk;
  const int z1 = min(a.iz1 + 1, z0 + a.zchunk);
  const int tx = (int)threadIdx.x % TX, ty = (int)threadIdx.x / TX;
  const int xt0 = a.ix0 + tix * TX, yt0 = a.iy0 + tiy * TY;
  const int x = xt0 + tx, y = yt0 + ty;
  // every thread inside the block stages its node (the halo of the box edge
  // rows); only threads inside the interior box compute
  const bool inblk = x < a.nx && y < a.ny;
  const bool active = x <= a.ix1 && y <= a.iy1;
  const long long P = a.P;
  const unsigned plane = (unsigned)a.nx * (unsigned)a.ny;
  const unsigned col = (unsigned)x + (unsigned)a.nx * (unsigned)y;

  // z queue: (v, fl_s) of levels z-3 .. z+3 (index k <-> level z-3+k); the
  // interior box keeps every level in [3, nz-4].  The first six iterations
  // (z0-6 .. z0-1) only fill the queue.
  T qv[7][3], qf[7][3];
#pragma unroll
  for (int k = 0; k < 7; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) { qv[k][c] = T(0); qf[k][c] = T(0); }

  for (int z = z0 - 6; z < z1; ++z) {
    const unsigned p = col + (unsigned)z * plane;
    // level z+3 of the own column enters the queue (loads issued here, the
    // operands formed after the LDS staging of level z)
    RawOps<T, MESH, S> rz;
    if (inblk) load_ops<T, MESH, 2>(a, p + 3u * plane, rz);
    if (z < z0) {
      if (inblk) make_ops<T, MESH, 2>(rz, qv[6], qf[6]);
#pragma unroll
      for (int k = 0; k < 6; ++k)
#pragma unroll
        for (int c = 0; c < 3; ++c) { qv[k][c] = qv[k + 1][c]; qf[k][c] = qf[k + 1][c]; }
      continue;
    }
    // MODE 3 finishes the velocity components of its node: their state, old
    // rate (and compensation) are loaded here, ahead of the LDS staging and
    // the differences that hide their latency
    S pre_f[3], pre_df[3], pre_c[3];
    if constexpr (MODE == 3) {
      if (active) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          pre_f[c] = ldp(a.Fin + c * P, p);
          pre_df[c] = ld_rate<NT>(a.A, atp(a.DF + c * P, p));
          if constexpr (std::is_same<S, float>::value) pre_c[c] = ld_stream<NT>(atp(a.Fc + c * P, p));
          else pre_c[c] = S(0);
        }
      }
    }
    // FP32 storage: the epilogue's material and geometry at the node
    // (lambda, mu, rho and the metric entries the staging below does not
    // load: X_m(2), or the Cartesian scales) are loaded here and the rest
    // kept from the staging, so no load of the epilogue waits behind the
    // barrier.  FP64 storage reloads them in the epilogue (the extra live
    // registers would spill).
    constexpr bool PRE = WQL_IPRE && sizeof(T) == 4;
    T em[3], eg[MESH ? 10 : 3];
    if (PRE && active) {
#pragma unroll
      for (int k = 0; k < 3; ++k) em[k] = ld_stream<NT>(atp(a.M + k * P, p));
      if constexpr (MESH) {
#pragma unroll
        for (int m = 0; m < 3; ++m) eg[3 * m + 2] = ldp(a.G + (3 * m + 2) * P, p);
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) eg[d] = ldp(a.G + d * P, p);
      }
    }
    // own node of level z: v from the queue, fl_q and fl_r
    if (inblk) {
      T vq[3], fq[3], fr[3];
      RawOps<T, MESH, S> r0, r1;
      load_ops<T, MESH, 0>(a, p, r0);
      load_ops<T, MESH, 1>(a, p, r1);
      make_ops<T, MESH, 0>(r0, vq, fq);
      make_ops<T, MESH, 1>(r1, vq, fr);
      if constexpr (PRE && MESH) {
        eg[9] = r0.g[0];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          eg[3 * m + 0] = r0.g[1 + m];
          eg[3 * m + 1] = r1.g[1 + m];
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        L[c][ty + 3][tx + 3] = qv[3][c];
        L[3 + c][ty + 3][tx + 3] = fq[c];
        L[6 + c][ty + 3][tx + 3] = fr[c];
      }
    }
    // halo: x strips (6 x TY) with fl_q, y strips (TX x 6) with fl_r
    constexpr int NHX = 6 * TY, NH = NHX + 6 * TX;
    for (int h = (int)threadIdx.x; h < NH; h += NTH) {
      int hx, hy;
      const bool xs = h < NHX;
      if (xs) {
        const int k = h % 6;
        hy = h / 6;
        hx = k < 3 ? k - 3 : TX + (k - 3);
      } else {
        const int k = (h - NHX) / TX;
        hx = (h - NHX) % TX;
        hy = k < 3 ? k - 3 : TY + (k - 3);
      }
      const int gx = xt0 + hx, gy = yt0 + hy;
      if (gx >= a.nx || gy >= a.ny) continue;   // gx, gy >= 3 by construction
      const unsigned gp = (unsigned)gx + (unsigned)a.nx * (unsigned)gy + (unsigned)z * plane;
      T hv[3], hf[3];
      if (xs) node_operands<T, MESH, 0>(a, gp, hv, hf);
      else node_operands<T, MESH, 1>(a, gp, hv, hf);
      const int fo = xs ? 3 : 6;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        L[c][hy + 3][hx + 3] = hv[c];
        L[fo + c][hy + 3][hx + 3] = hf[c];
      }
    }
    __syncthreads();
    // level z+3 joins the queue (its loads were issued at the top)
    if (inblk) make_ops<T, MESH, 2>(rz, qv[6], qf[6]);

    if (active) {
      // central-stencil differences (JU_xJU_yJU_z6.f90:53-59), unscaled
      T Dv[3][3], Df[3][3];
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int c = 0; c < 3; ++c) { Dv[d][c] = T(0); Df[d][c] = T(0); }
      // one LDS operand plane at a time; pin() makes each sum complete where
      // it is written, so the compiler neither sinks the sums towards their
      // use nor keeps all 72 LDS values live at once
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        for_taps_interior([&](int off, double cd) {
          const T ck = (T)cd;
          Dv[0][c] += ck * L[c][ty + 3][tx + 3 + off];
          Dv[1][c] += ck * L[c][ty + 3 + off][tx + 3];
        });
        pin(Dv[0][c]);
        pin(Dv[1][c]);
        for_taps_interior([&](int off, double cd) {
          const T ck = (T)cd;
          Df[0][c] += ck * L[3 + c][ty + 3][tx + 3 + off];
          Df[1][c] += ck * L[6 + c][ty + 3 + off][tx + 3];
        });
        pin(Df[0][c]);
        pin(Df[1][c]);
      }
      for_taps_interior([&](int off, double cd) {
        const T ck = (T)cd;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          Dv[2][c] += ck * qv[3 + off][c];
          Df[2][c] += ck * qf[3 + off][c];
        }
      });
      const unsigned idu = p;
      const T lam = PRE ? em[0] : ld_stream<NT>(atp(a.M, idu));
      const T mu = PRE ? em[1] : ld_stream<NT>(atp(a.M + P, idu));
      const T rho = PRE ? em[2] : ld_stream<NT>(atp(a.M + 2 * P, idu));
#define WQL_EG(k) (PRE ? eg[k] : ldp(a.G + (k) * P, idu))
      const T l2m = lam + T(2) * mu;
      T gr[3][3];   // gr[m][c] = dv_c/dx_m
      T DU[9];
      if constexpr (MESH) {
        const T ihd[3] = {(T)a.ih[0], (T)a.ih[1], (T)a.ih[2]};
        const T J = WQL_EG(9);
        const T rhoJ_inv = T(1) / (rho * J);
#pragma unroll
        for (int c = 0; c < 3; ++c)
          DU[c] = (ihd[0] * Df[0][c] + ihd[1] * Df[1][c] + ihd[2] * Df[2][c]) * rhoJ_inv;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          const T X0 = WQL_EG(3 * m + 0) * ihd[0];
          const T X1 = WQL_EG(3 * m + 1) * ihd[1];
          const T X2 = WQL_EG(3 * m + 2) * ihd[2];
#pragma unroll
          for (int c = 0; c < 3; ++c) gr[m][c] = X0 * Dv[0][c] + X1 * Dv[1][c] + X2 * Dv[2][c];
        }
      } else {
        T sd[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) sd[d] = WQL_EG(d) * (T)a.ih[d];
#undef WQL_EG
        const T rho_inv = T(1) / rho;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          DU[c] = (sd[0] * Df[0][c] + sd[1] * Df[1][c] + sd[2] * Df[2][c]) * rho_inv;
#pragma unroll
        for (int m = 0; m < 3; ++m)
#pragma unroll
          for (int c = 0; c < 3; ++c) gr[m][c] = sd[m] * Dv[m][c];
      }
      DU[3] = l2m * gr[0][0] + lam * (gr[1][1] + gr[2][2]);
      DU[4] = l2m * gr[1][1] + lam * (gr[0][0] + gr[2][2]);
      DU[5] = l2m * gr[2][2] + lam * (gr[0][0] + gr[1][1]);
      DU[6] = mu * (gr[1][0] + gr[0][1]);
      DU[7] = mu * (gr[2][0] + gr[0][2]);
      DU[8] = mu * (gr[2][1] + gr[1][2]);
      double srcv[6];
      const bool has_src = a.nsrc > 0 && source_at<T, MESH>(a, x, y, z, (long long)idu, srcv);
      if constexpr (MODE == 3) {
        // velocity RK update here, stress rates and strain rates to W for
        // k_memory_update (as k_node MODE 3)
        const T e[6] = {gr[0][0], gr[1][1], gr[2][2], gr[1][0] + gr[0][1], gr[2][0] + gr[0][2],
                        gr[2][1] + gr[1][2]};
#pragma unroll
        for (int c = 0; c < 3; ++c)
          rk_finish_pre<NT>(a, c, idu, scale_rate(a.A, pre_df[c]) + (double)DU[c], pre_f[c],
                            pre_c[c]);
        // stress rates only where the source adds to them: elsewhere in the
        // strict interior they are Hooke's law of e (hooke_node)
        if (has_src) {
#pragma unroll
          for (int c = 3; c < 9; ++c)
            st_stream<NT>(atp(a.W + (c - 3) * P, idu), DU[c] - (T)srcv[c - 3]);
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) st_stream<NT>(atp(a.W + (6 + c) * P, idu), e[c]);
      } else {
        if constexpr (NMECH > 0) {
          const T e[6] = {gr[0][0], gr[1][1], gr[2][2], gr[1][0] + gr[0][1],
                          gr[2][0] + gr[0][2], gr[2][1] + gr[1][2]};
          memory_update_node<T, NMECH, NT>(a, idu, mu, l2m, e, DU);
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) {
          double df = scale_rate(a.A, ld_rate<NT>(a.A, atp(a.DF + c * P, idu))) + (double)DU[c];
          if (has_src && c >= 3) df -= srcv[c - 3];
          rk_finish<NT>(a, c, idu, df);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) { qv[k][c] = qv[k + 1][c]; qf[k][c] = qf[k + 1][c]; }
  }
}

// SAT on one face (impose_boundary_condition, unfused): DF -= tau0/h*structure(ubc),
// PML DQ -= tau0/h*ubc
__global__ __launch_bounds__(256) void k_sat_face(const StepArgs<double> a, int f) {
  const int axis = f >> 1;
  const int n1 = axis == 0 ? a.ny : a.nx;
  const int n2 = axis == 2 ? a.ny : a.nz;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)n1 * n2) return;
  const int j = (int)(t / n1), i = (int)(t - (long long)j * n1);
  const int fixed = (f & 1) ? (axis == 0 ? a.nx - 1 : (axis == 1 ? a.ny - 1 : a.nz - 1)) : 0;
  int x, y, z;
  if (axis == 0) { x = fixed; y = i; z = j; }
  else if (axis == 1) { x = i; y = fixed; z = j; }
  else { x = i; y = j; z = fixed; }
  const long long P = a.P;
  const long long id = x + (long long)a.nx * (y + (long long)a.ny * z);
  const double lam = a.M[id], mu = a.M[P + id], rho = a.M[2 * P + id];
  double u[9], ubc[9], Ux[9];
  for (int c = 0; c < 9; ++c) u[c] = a.Fin[c * P + id];
  sat_rate(a, f, t, (long long)n1 * n2, u, lam, mu, rho, ubc, Ux);
  const double pen = a.pen[axis];
  for (int c = 0; c < 9; ++c) a.DF[c * P + id] = a.DF[c * P + id] - pen * Ux[c];
  if (a.pml[f]) {
    long long si;
    double d;
    if (pml_slab(a, f, x, y, z, &si, &d)) {
      const long long Ps = slab_plane(a, f);
      for (int c = 0; c < 9; ++c) a.pDQ[f][c * Ps + si] = a.pDQ[f][c * Ps + si] - pen * ubc[c];
    }
  }
}

// set_moment_tensor over the +-4 (order 4) or +-2 (order 2) node box
__global__ __launch_bounds__(256) void k_source(const StepArgs<double> a, int s, int mesh) {
  const DevSource& S = a.src[s];
  const int r = source_radius(S);
  const int w = 2 * r + 1;
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= w * w * w) return;
  const int x = S.near[0] - r + t % w;
  const int y = S.near[1] - r + (t / w) % w;
  const int z = S.near[2] - r + t / (w * w);
  if (x < 0 || x >= a.nx || y < 0 || y >= a.ny || z < 0 || z >= a.nz) return;
  const long long P = a.P;
  const long long id = x + (long long)a.nx * (y + (long long)a.ny * z);
  double delta;
  if (S.order == 0) {
    delta = mollified_delta(a, x, y, z, id, S);
  } else {
    const double qx = a.G[id];
    const double ry = a.G[(mesh ? 4 : 1) * P + id];
    const double sz = a.G[(mesh ? 8 : 2) * P + id];
    const double hx = a.h[0] / qx, hy = a.h[1] / ry, hz = a.h[2] / sz;
    const double dx = singular_source(x, S.near[0], hx, S.off[0] / hx, S.order, S.alpha);
    const double dy = singular_source(y, S.near[1], hy, S.off[1] / hy, S.order, S.alpha);
    const double dz = singular_source(z, S.near[2], hz, S.off[2] / hz, S.order, S.alpha);
    delta = dx * dy * dz;
  }
  for (int c = 0; c < 6; ++c) a.DF[(3 + c) * P + id] = a.DF[(3 + c) * P + id] - S.am[c] * delta;
}

// rk_vector kernels (rk_vector_openacc.f90:22-75): no FMA contraction, so the
// result is bit-identical to the CPU expressions.
template <typename T>
__global__ __launch_bounds__(256) void k_scale(T* x, long long n, double A) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  x[i] = scale_rate(A, x[i]);
}
__device__ __forceinline__ double mul_add_rn(double s, double inc, double r) {
  return __dadd_rn(s, __dmul_rn(inc, r));
}
__device__ __forceinline__ float mul_add_rn(float s, double inc, float r) {
  return __fadd_rn(s, __fmul_rn((float)inc, r));
}
template <typename T>
__global__ __launch_bounds__(256) void k_update(T* x, const T* r, long long n, double inc) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  x[i] = mul_add_rn(x[i], inc, r[i]);
}
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* x, long long n, double v) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  x[i] = (T)v;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gather(T* dst, const double* src, int nx, int ny, int nz,
                                                long long sx, long long sy, long long sz,
                                                long long sc, long long off0, int c) {
  const long long plane = (long long)nx * ny;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= plane) return;
  const int z = blockIdx.y;
  const int y = (int)(t / nx), x = (int)(t - (long long)y * nx);
  const double v = src[off0 + x * sx + y * sy + z * sz + c * sc];
  dst[(long long)c * plane * nz + t + z * plane] = (T)v;
}
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(double* dst, const T* src, int nx, int ny, int nz,
                                                 long long sx, long long sy, long long sz,
                                                 long long sc, long long off0, int c) {
  const long long plane = (long long)nx * ny;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= plane) return;
  const int z = blockIdx.y;
  const int y = (int)(t / nx), x = (int)(t - (long long)y * nx);
  dst[off0 + x * sx + y * sy + z * sz + c * sc] =
      (double)src[(long long)c * plane * nz + t + z * plane];
}

template <typename S>
__global__ void k_sample(const S* F, long long P, const long long* idx, int n, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9 * n) return;
  const int r = t / 9, c = t - 9 * (t / 9);
  out[t] = F[c * P + idx[r]];
}

// derivatives() on the device (grid.f90:398-471 with metrics.f90 D6x/D6y/D6z)
__global__ __launch_bounds__(256) void k_metrics(int nx, int ny, int nz, const double* X,
                                                 double* mx, double* my, double* mz, double* J) {
  const long long plane = (long long)nx * ny;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= plane) return;
  const int z = blockIdx.y;
  const int y = (int)(t / nx), x = (int)(t - (long long)y * nx);
  const long long P = plane * nz;
  const long long id = t + z * plane;
  const long long st[3] = {1, nx, plane};
  const int pos[3] = {x, y, z}, n[3] = {nx, ny, nz};
  double D[3][3];
  for (int d = 0; d < 3; ++d) {
    const double ih = 1.0 / (1.0 / (double)(n[d] - 1));
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for_taps(pos[d], n[d], [&](int off, double c) {
      const long long p = id + (long long)off * st[d];
      s0 += c * X[p];
      s1 += c * X[P + p];
      s2 += c * X[2 * P + p];
    });
    D[d][0] = ih * s0;
    D[d][1] = ih * s1;
    D[d][2] = ih * s2;
  }
  const double* Xq = D[0];
  const double* Xr = D[1];
  const double* Xs = D[2];
  const double j = Xq[0] * (Xr[1] * Xs[2] - Xr[2] * Xs[1]) - Xq[1] * (Xr[0] * Xs[2] - Xr[2] * Xs[0]) +
                   Xq[2] * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
  const double ij = 1.0 / j;
  J[id] = j;
  mx[id] = ij * (Xr[1] * Xs[2] - Xr[2] * Xs[1]);
  mx[P + id] = ij * (Xq[2] * Xs[1] - Xq[1] * Xs[2]);
  mx[2 * P + id] = ij * (Xq[1] * Xr[2] - Xq[2] * Xr[1]);
  my[id] = ij * (Xr[2] * Xs[0] - Xr[0] * Xs[2]);
  my[P + id] = ij * (Xq[0] * Xs[2] - Xq[2] * Xs[0]);
  my[2 * P + id] = ij * (Xq[2] * Xr[0] - Xq[0] * Xr[2]);
  mz[id] = ij * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
  mz[P + id] = ij * (Xq[1] * Xs[0] - Xq[0] * Xs[1]);
  mz[2 * P + id] = ij * (Xq[0] * Xr[1] - Xq[1] * Xr[0]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline dim3 plane_grid(int nx, int ny, int nz) {
  const long long plane = (long long)nx * ny;
  return dim3((unsigned)((plane + 255) / 256), (unsigned)nz, 1);
}

template <typename T, int MODE, bool NT, typename S>
static hipError_t launch_node_nt(const StepArgs<T, S>& a, int mesh, hipStream_t s) {
  dim3 g;
  if (a.nbox > 0) {
    g = dim3((unsigned)((a.box_start[a.nbox] + 255) / 256));
  } else {
    const long long nbands = (a.ny + a.band_rows - 1) / a.band_rows;
    g = dim3((unsigned)(nbands * a.nz * a.segs));
  }
  const dim3 b(256);
#define WQL_CASE(MS, NM)                                                   \
  if (mesh == MS && a.nmech == NM) {                                       \
    hipLaunchKernelGGL((k_node<T, MS, NM, MODE, NT, S>), g, b, 0, s, a);   \
    return hipGetLastError();                                              \
  }
  WQL_CASE(0, 0) WQL_CASE(0, 4) WQL_CASE(0, 8)
  WQL_CASE(1, 0) WQL_CASE(1, 4) WQL_CASE(1, 8)
#undef WQL_CASE
  return hipErrorInvalidValue;
}

template <typename T, int MODE, typename S>
static hipError_t launch_node(const StepArgs<T, S>& a, int mesh, hipStream_t s) {
  if ((MODE == 0 || MODE == 3) && a.nontemporal) return launch_node_nt<T, MODE, true>(a, mesh, s);
  return launch_node_nt<T, MODE, false>(a, mesh, s);
}

template <typename T, int NM, int NP, bool NT, typename S>
static void launch_memory_update_n(const StepArgs<T, S>& a, dim3 g, hipStream_t s) {
  const bool rd = a.A != 0.0, wr = !a.dead_rates;
  if (rd && wr) hipLaunchKernelGGL((k_memory_update_n<T, NM, NP, NT, S, true, true>), g, dim3(256), 0, s, a);
  else if (rd) hipLaunchKernelGGL((k_memory_update_n<T, NM, NP, NT, S, true, false>), g, dim3(256), 0, s, a);
  else if (wr) hipLaunchKernelGGL((k_memory_update_n<T, NM, NP, NT, S, false, true>), g, dim3(256), 0, s, a);
  else hipLaunchKernelGGL((k_memory_update_n<T, NM, NP, NT, S, false, false>), g, dim3(256), 0, s, a);
}

template <typename T, bool NT, typename S>
static hipError_t launch_memory_update(const StepArgs<T, S>& a, hipStream_t s) {
  if (a.node1 <= a.node0) return hipSuccess;
  if (a.mem_npt == 4 && a.node0 % 4 == 0 && a.node1 % 4 == 0) {
    const dim3 g((unsigned)((a.node1 - a.node0 + 1023) / 1024));
    if (a.nmech == 4) launch_memory_update_n<T, 4, 4, NT>(a, g, s);
    else if (a.nmech == 8) launch_memory_update_n<T, 8, 4, NT>(a, g, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
  }
  if (a.mem_npt >= 2 && a.node0 % 2 == 0 && a.node1 % 2 == 0) {
    const dim3 g((unsigned)((a.node1 - a.node0 + 511) / 512));
    if (a.nmech == 4) launch_memory_update_n<T, 4, 2, NT>(a, g, s);
    else if (a.nmech == 8) launch_memory_update_n<T, 8, 2, NT>(a, g, s);
    else return hipErrorInvalidValue;
    return hipGetLastError();
  }
  const dim3 g((unsigned)((a.node1 - a.node0 + 255) / 256));
  if (a.nmech == 4) hipLaunchKernelGGL((k_memory_update<T, 4, NT, S>), g, dim3(256), 0, s, a);
  else if (a.nmech == 8) hipLaunchKernelGGL((k_memory_update<T, 8, NT, S>), g, dim3(256), 0, s, a);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename T, int TX, int TY, typename S>
static hipError_t launch_interior_tile(const StepArgs<T, S>& a, int mesh, bool to_w, hipStream_t s) {
  const long long bx = a.ix1 - a.ix0 + 1, by = a.iy1 - a.iy0 + 1, bz = a.iz1 - a.iz0 + 1;
  const long long ntx = (bx + TX - 1) / TX, nty = (by + TY - 1) / TY;
  const long long nzc = (bz + a.zchunk - 1) / a.zchunk;
  const dim3 g((unsigned)(ntx * nty * nzc)), b(TX * TY);
#define WQL_CASE(MS, NM, MD)                                                        \
  if (mesh == MS && a.nmech == NM && (MD == 3) == to_w) {                           \
    hipLaunchKernelGGL((k_interior<T, MS, NM, MD, false, TX, TY, S>), g, b, 0, s, a);  \
    return hipGetLastError();                                                       \
  }
  WQL_CASE(0, 0, 0) WQL_CASE(0, 4, 3) WQL_CASE(0, 8, 3) WQL_CASE(0, 4, 0) WQL_CASE(0, 8, 0)
  WQL_CASE(1, 0, 0) WQL_CASE(1, 4, 3) WQL_CASE(1, 8, 3) WQL_CASE(1, 4, 0) WQL_CASE(1, 8, 0)
#undef WQL_CASE
  return hipErrorInvalidValue;
}

// tile shape of the interior kernel: zm = 1 -> 64x8, 2 -> 32x16, 3 -> 64x4
template <typename T, typename S>
static hipError_t launch_interior(const StepArgs<T, S>& a, int mesh, bool to_w, hipStream_t s) {
  switch (a.zm) {
    case 2: return launch_interior_tile<T, 32, 16>(a, mesh, to_w, s);
    case 3: return launch_interior_tile<T, 64, 4>(a, mesh, to_w, s);
    default: return launch_interior_tile<T, 64, 8>(a, mesh, to_w, s);
  }
}

// The six boxes of the block outside the strict interior box, full x rows
// wherever possible: z slabs, then y slabs of the remaining z range, then x
// slabs of the remaining (y, z) range.
template <typename T, typename S>
static void set_shell_boxes(StepArgs<T, S>& a) {
  int nb = 0;
  auto add = [&](int x0, int x1, int y0, int y1, int z0, int z1) {
    if (x0 > x1 || y0 > y1 || z0 > z1) return;
    const int v[6] = {x0, x1, y0, y1, z0, z1};
    for (int k = 0; k < 6; ++k) a.box[nb][k] = v[k];
    a.box_start[nb + 1] =
        a.box_start[nb] + (long long)(x1 - x0 + 1) * (y1 - y0 + 1) * (long long)(z1 - z0 + 1);
    ++nb;
  };
  a.box_start[0] = 0;
  add(0, a.nx - 1, 0, a.ny - 1, 0, a.iz0 - 1);
  add(0, a.nx - 1, 0, a.ny - 1, a.iz1 + 1, a.nz - 1);
  add(0, a.nx - 1, 0, a.iy0 - 1, a.iz0, a.iz1);
  add(0, a.nx - 1, a.iy1 + 1, a.ny - 1, a.iz0, a.iz1);
  add(0, a.ix0 - 1, a.iy0, a.iy1, a.iz0, a.iz1);
  add(a.ix1 + 1, a.nx - 1, a.iy0, a.iy1, a.iz0, a.iz1);
  a.nbox = nb;
}

template <typename T, typename S>
static bool interior_on(const StepArgs<T, S>& a) {
  return a.zm > 0 && a.zchunk > 0 && a.ix0 <= a.ix1 && a.iy0 <= a.iy1 && a.iz0 <= a.iz1;
}

// anelastic stages with the W hand-off: node kernels write rates and strain
// rates, k_memory_update finishes every node
template <typename T, typename S>
static bool split_stage(const StepArgs<T, S>& a) {
  return a.nmech > 0 && a.split && a.W;
}

template <typename T, typename S>
hipError_t launch_stage_part(const StepArgs<T, S>& a, int mesh, StagePart part, hipStream_t s) {
  switch (part) {
    case StagePart::kInterior:
      return interior_on(a) ? launch_interior<T>(a, mesh, split_stage(a), s) : hipSuccess;
    case StagePart::kShell: {
      StepArgs<T, S> sh = a;
      if (interior_on(a)) set_shell_boxes(sh);
      return split_stage(a) ? launch_node<T, 3>(sh, mesh, s) : launch_node<T, 0>(sh, mesh, s);
    }
    default:
      if (!split_stage(a)) return hipSuccess;
      return (a.nontemporal || a.mem_nt) ? launch_memory_update<T, true>(a, s)
                           : launch_memory_update<T, false>(a, s);
  }
}

template <typename T, typename S>
bool stage_has_memory_part(const StepArgs<T, S>& a) {
  return split_stage(a);
}

// One part of the stage over the z levels [z0, z1): the interior box and the
// shell boxes cut to those levels, or the flat node range of those levels
// (the last range also covers the plane padding).  The stencil parts read F
// of levels z0-3 .. z1+2; every part writes only nodes of its levels.
template <typename T, typename S>
hipError_t launch_stage_part_z(const StepArgs<T, S>& a, int mesh, StagePart part, int z0, int z1,
                               hipStream_t s) {
  if (z0 >= z1) return hipSuccess;
  switch (part) {
    case StagePart::kInterior: {
      if (!interior_on(a)) return hipSuccess;
      StepArgs<T, S> b = a;
      b.iz0 = a.iz0 > z0 ? a.iz0 : z0;
      b.iz1 = a.iz1 < z1 - 1 ? a.iz1 : z1 - 1;
      if (b.iz0 > b.iz1) return hipSuccess;
      return launch_interior<T>(b, mesh, split_stage(a), s);
    }
    case StagePart::kShell: {
      StepArgs<T, S> sh = a;
      if (interior_on(a)) {
        set_shell_boxes(sh);
      } else {
        sh.nbox = 1;
        const int v[6] = {0, a.nx - 1, 0, a.ny - 1, 0, a.nz - 1};
        for (int k = 0; k < 6; ++k) sh.box[0][k] = v[k];
      }
      // cut every box to [z0, z1 - 1]
      StepArgs<T, S> cut = sh;
      int nb = 0;
      cut.box_start[0] = 0;
      for (int i = 0; i < sh.nbox; ++i) {
        int bz0 = sh.box[i][4] > z0 ? sh.box[i][4] : z0;
        int bz1 = sh.box[i][5] < z1 - 1 ? sh.box[i][5] : z1 - 1;
        if (bz0 > bz1) continue;
        for (int k = 0; k < 4; ++k) cut.box[nb][k] = sh.box[i][k];
        cut.box[nb][4] = bz0;
        cut.box[nb][5] = bz1;
        cut.box_start[nb + 1] = cut.box_start[nb] +
                                (long long)(sh.box[i][1] - sh.box[i][0] + 1) *
                                    (sh.box[i][3] - sh.box[i][2] + 1) * (long long)(bz1 - bz0 + 1);
        ++nb;
      }
      if (nb == 0) return hipSuccess;
      cut.nbox = nb;
      return split_stage(a) ? launch_node<T, 3>(cut, mesh, s) : launch_node<T, 0>(cut, mesh, s);
    }
    default: {
      if (!split_stage(a)) return hipSuccess;
      StepArgs<T, S> b = a;
      const long long plane = (long long)a.nx * a.ny;
      b.node0 = z0 * plane;
      b.node1 = z1 >= a.nz ? a.P : z1 * plane;
      return (a.nontemporal || a.mem_nt) ? launch_memory_update<T, true>(b, s)
                           : launch_memory_update<T, false>(b, s);
    }
  }
}

template <typename T, typename S>
hipError_t launch_stage_fused(const StepArgs<T, S>& a, int mesh, hipStream_t s) {
  for (StagePart p : {StagePart::kInterior, StagePart::kShell, StagePart::kMemory}) {
    hipError_t e = launch_stage_part<T>(a, mesh, p, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
template hipError_t launch_stage_part<double>(const StepArgs<double>&, int, StagePart, hipStream_t);
template hipError_t launch_stage_part<float>(const StepArgs<float>&, int, StagePart, hipStream_t);
template hipError_t launch_stage_fused<double>(const StepArgs<double>&, int, hipStream_t);
template hipError_t launch_stage_fused<float>(const StepArgs<float>&, int, hipStream_t);
template hipError_t launch_stage_part<float, float>(const StepArgs<float, float>&, int, StagePart,
                                                    hipStream_t);
template hipError_t launch_stage_fused<float, float>(const StepArgs<float, float>&, int, hipStream_t);
template hipError_t launch_stage_part_z<double>(const StepArgs<double>&, int, StagePart, int, int,
                                                hipStream_t);
template hipError_t launch_stage_part_z<float>(const StepArgs<float>&, int, StagePart, int, int,
                                               hipStream_t);
template hipError_t launch_stage_part_z<float, float>(const StepArgs<float, float>&, int, StagePart,
                                                      int, int, hipStream_t);
template bool stage_has_memory_part<double>(const StepArgs<double>&);
template bool stage_has_memory_part<float>(const StepArgs<float>&);
template bool stage_has_memory_part<float, float>(const StepArgs<float, float>&);

hipError_t launch_rhs(const StepArgs<double>& a, int mesh, RhsPart part, hipStream_t s) {
  return part == RhsPart::kAll ? launch_node<double, 1>(a, mesh, s)
                               : launch_node<double, 2>(a, mesh, s);
}

hipError_t launch_sat(const StepArgs<double>& a, hipStream_t s) {
  for (int f = 0; f < 6; ++f) {
    if (a.bc[f] <= 0) continue;
    const int axis = f >> 1;
    const long long n1 = axis == 0 ? a.ny : a.nx;
    const long long n2 = axis == 2 ? a.ny : a.nz;
    hipLaunchKernelGGL(k_sat_face, dim3((unsigned)((n1 * n2 + 255) / 256)), dim3(256), 0, s, a, f);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_source(const StepArgs<double>& a, int mesh, hipStream_t s) {
  for (int k = 0; k < a.nsrc; ++k) {
    const int r = source_radius(a.src[k]);
    const int w = 2 * r + 1;
    hipLaunchKernelGGL(k_source, dim3((w * w * w + 255) / 256), dim3(256), 0, s, a, k, mesh);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename T>
hipError_t launch_scale(T* x, long long n, double A, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, A);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_update(T* x, const T* r, long long n, double inc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_update<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, r, n, inc);
  return hipGetLastError();
}
template <typename T>
hipError_t launch_fill(T* x, long long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n, v);
  return hipGetLastError();
}
template hipError_t launch_scale<double>(double*, long long, double, hipStream_t);
template hipError_t launch_scale<float>(float*, long long, double, hipStream_t);
template hipError_t launch_update<double>(double*, const double*, long long, double, hipStream_t);
template hipError_t launch_update<float>(float*, const float*, long long, double, hipStream_t);
template hipError_t launch_fill<double>(double*, long long, double, hipStream_t);
template hipError_t launch_fill<float>(float*, long long, double, hipStream_t);

// dst = (float)src, round to nearest: the float copy of the RK state
__global__ __launch_bounds__(256) void k_round_float(float* dst, const double* src, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  dst[i] = (float)src[i];
}
hipError_t launch_round_float(float* dst, const double* src, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_round_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gather(T* dst, const double* src, int nx, int ny, int nz, int ncomp,
                         long long sx, long long sy, long long sz, long long sc, long long off0,
                         hipStream_t s) {
  for (int c = 0; c < ncomp; ++c) {
    hipLaunchKernelGGL(k_gather<T>, plane_grid(nx, ny, nz), dim3(256), 0, s, dst, src, nx, ny, nz,
                       sx, sy, sz, sc, off0, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
template <typename T>
hipError_t launch_scatter(double* dst, const T* src, int nx, int ny, int nz, int ncomp,
                          long long sx, long long sy, long long sz, long long sc, long long off0,
                          hipStream_t s) {
  for (int c = 0; c < ncomp; ++c) {
    hipLaunchKernelGGL(k_scatter<T>, plane_grid(nx, ny, nz), dim3(256), 0, s, dst, src, nx, ny, nz,
                       sx, sy, sz, sc, off0, c);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
template hipError_t launch_gather<double>(double*, const double*, int, int, int, int, long long,
                                          long long, long long, long long, long long, hipStream_t);
template hipError_t launch_gather<float>(float*, const double*, int, int, int, int, long long,
                                         long long, long long, long long, long long, hipStream_t);
template hipError_t launch_scatter<double>(double*, const double*, int, int, int, int, long long,
                                           long long, long long, long long, long long, hipStream_t);
template hipError_t launch_scatter<float>(double*, const float*, int, int, int, int, long long,
                                          long long, long long, long long, long long, hipStream_t);

template <typename S>
hipError_t launch_sample(const S* F, long long P, const long long* idx, int n, double* out,
                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample<S>, dim3((9 * n + 255) / 256), dim3(256), 0, s, F, P, idx, n, out);
  return hipGetLastError();
}
template hipError_t launch_sample<double>(const double*, long long, const long long*, int, double*,
                                          hipStream_t);
template hipError_t launch_sample<float>(const float*, long long, const long long*, int, double*,
                                         hipStream_t);

hipError_t launch_metrics(int nx, int ny, int nz, const double* X, double* mx, double* my,
                          double* mz, double* J, hipStream_t s) {
  hipLaunchKernelGGL(k_metrics, plane_grid(nx, ny, nz), dim3(256), 0, s, nx, ny, nz, X, mx, my, mz, J);
  return hipGetLastError();
}

}  // namespace wql
