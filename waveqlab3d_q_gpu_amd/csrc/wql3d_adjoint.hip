// Discrete adjoint of the fused RK stage (HIP, gfx950).
//
// Forward stage s of the 2N-storage scheme (time_step.f90:109-147), with the
// state u = (F, eta) and the rate registers r = (DF, Deta):
//   r <- A_s r + L(u; m) ;  u <- u + B_s dt r
// L(u; m) = K(m) u + f(t): K is linear in u (SBP derivatives, Hooke, memory
// variables, SAT penalties), f the moment-tensor source, m = (rho, lambda, mu)
// per node.  Reverse of stage s, given the adjoints (ub, rb) of its outputs:
//   rb += B_s dt ub            (from u <- u + B dt r)
//   g  += (dL/dm)^T rb  at the stage's input state u
//   ub += K^T rb
//   rb  = A_s rb
// K acts pointwise except for the derivatives, and every material term of L
// at a node uses only that node's material, so dL/dm is local.  A reverse
// stage is four kernels: k_adj_grad evaluates the forward derivatives of the
// stage input at each node and contracts the local material Jacobian with
// rb + B dt ub (formed on the fly); k_adj_pre stores rb + B dt ub, forms the
// pointwise adjoint fields Z of K^T = (pointwise)^T D^T (pointwise)^T and the
// memory-variable coupling, and leaves A rb in the rate adjoints; k_adj_sat
// adds the SAT penalties' transpose and material terms on the boundary
// faces (their dual-number arithmetic would cost the per-node kernels half
// their waves); k_adj_stencil applies the transposed SBP operators D_d^T to
// Z.  The two streaming updates of rb thus ride on the node kernels' own
// reads and writes of rb instead of four passes of their own.
//
// Flux form of the forward operator (as in the interior kernel):
//   rho J dv_c/dt = sum_d D_d(fl_d[c]) / h_d,  fl_d[c] = J sum_m X_m(d) sigma(m,c)
//   dv_c/dx_m     = sum_d X_m(d) D_d(v_c) / h_d
// (Cartesian: X diagonal, J = 1, and the metric factor G_d/h_d multiplies the
// derivative at the output node.)  PML slabs are not differentiated.
#include "wql3d_device.h"
#include "wql3d_adjoint.h"

namespace wql {

// Entry D[i][j] of the unscaled SBP-6 first-derivative matrix on n points.
__device__ __forceinline__ double sbp_entry(int i, int j, int n) {
  if (i < 6) return (j >= 0 && j < 9) ? kSbpL[i][j] : 0.0;
  if (i >= n - 6) {
    const int r = n - 1 - i, k = j - (n - 9);
    return (k >= 0 && k < 9) ? -kSbpL[r][8 - k] : 0.0;
  }
  switch (j - i) {
    case -3: return -1.0 / 60.0;
    case -2: return 0.15;
    case -1: return -0.75;
    case 1: return 0.75;
    case 2: return -0.15;
    case 3: return 1.0 / 60.0;
    default: return 0.0;
  }
}

// Apply f(offset, D[j+offset][j]) for every non-zero entry of column j of D:
// (D^T p)_j = sum f over p at node j+offset.  Columns 9..n-10 only meet
// central rows, where D[j-k][j] = c_k (the central coefficients).
template <typename Fn>
__device__ __forceinline__ void for_taps_T(int j, int n, Fn&& f) {
  if (j >= 9 && j < n - 9) {
    f(-3, 1.0 / 60.0);
    f(-2, -0.15);
    f(-1, 0.75);
    f(1, -0.75);
    f(2, 0.15);
    f(3, -1.0 / 60.0);
  } else {
    for (int o = -8; o <= 8; ++o) {
      const int i = j + o;
      if (i < 0 || i >= n) continue;
      const double c = sbp_entry(i, j, n);
      if (c != 0.0) f(o, c);
    }
  }
}

// Dual number with derivatives along (rho, lambda, mu).
struct D3 {
  double v, d[3];
  __device__ D3() : v(0.0), d{0.0, 0.0, 0.0} {}
  __device__ explicit D3(double x) : v(x), d{0.0, 0.0, 0.0} {}
};
__device__ __forceinline__ D3 operator+(const D3& a, const D3& b) {
  D3 r(a.v + b.v);
  for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
__device__ __forceinline__ D3 operator-(const D3& a, const D3& b) {
  D3 r(a.v - b.v);
  for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
__device__ __forceinline__ D3 operator*(const D3& a, const D3& b) {
  D3 r(a.v * b.v);
  for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
__device__ __forceinline__ D3 operator/(const D3& a, const D3& b) {
  D3 r(a.v / b.v);
  for (int k = 0; k < 3; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
template <typename R>
__device__ __forceinline__ D3 operator*(R a, const D3& b) { return D3((double)a) * b; }
template <typename R>
__device__ __forceinline__ D3 operator*(const D3& a, R b) { return a * D3((double)b); }
__device__ __forceinline__ D3 sqrt(const D3& a) {
  D3 r(::sqrt(a.v));
  for (int k = 0; k < 3; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
  return r;
}

// node (x, y, z) of flat index id < nx*ny*nz
template <typename T, typename St>
__device__ __forceinline__ bool flat_node(const StepArgs<T, St>& a, long long id, int& x, int& y, int& z) {
  const long long nodes = (long long)a.nx * a.ny * a.nz;
  if (id >= nodes) return false;
  const long long plane = (long long)a.nx * a.ny;
  z = (int)(id / plane);
  const long long r = id - (long long)z * plane;
  y = (int)(r / a.nx);
  x = (int)(r - (long long)y * a.nx);
  return true;
}

// SAT rate vector of face f at a node for state u: -pen_f * structure(ubc(u))
template <typename S, typename T, typename St>
__device__ __forceinline__ void sat_rate(const StepArgs<T, St>& a, int f, long long fidx, long long NF,
                                         const S u[9], S lam, S mu, S rho, S out[9]) {
  T fr[10];
  face_frame(a, f, fidx, NF, fr);
  S ubc[9], Ux[9];
  sat_ubc_frame<S, T>(fr, a.bc[f], f, u, lam, mu, rho, ubc);
  sat_structure<S>(ubc, lam, mu, rho, Ux);
  const S pen((double)(T)a.pen[f >> 1]);
#pragma unroll
  for (int c = 0; c < 9; ++c) out[c] = S(0.0) - pen * Ux[c];
}

// rb + Bdt ub in the order of k_update (the rk_vector update of the
// reference order path), so the fused reverse stage equals the one built
// from separate update / scale launches bit for bit
__device__ __forceinline__ double adj_inc(double r, double inc, double u) {
  return __dadd_rn(r, __dmul_rn(inc, u));
}
__device__ __forceinline__ float adj_inc(float r, double inc, float u) {
  return __fadd_rn(r, __fmul_rn((float)inc, u));
}

// ---------------------------------------------------------------------------
// gradient, interior part: g += (dL/dm)^T rb at the stage input state (its F,
// a.Fin) for the terms that need the state's derivatives (velocity rates,
// Hooke, memory variables); rb = DFb + Bdt Fb and rb_eta = DEb + Bdt Eb are
// formed here without being stored (k_adj_pre stores them)
// ---------------------------------------------------------------------------
template <typename T, typename St, int MESH, int NMECH>
__global__ __launch_bounds__(256) void k_adj_grad(const StepArgs<T, St> a, const AdjArgs<T> j,
                                                  double Bdt) {
  int x, y, z;
  const long long id = (long long)xcd_logical(blockIdx.x, gridDim.x, j.xcd) * 256 + threadIdx.x;
  if (!flat_node(a, id, x, y, z)) return;
  const long long P = a.P;
  const int pos[3] = {x, y, z};
  T U[3][9];
  if (MESH) derivs_curv<T>(a, pos, id, U);
  else derivs_cart<T>(a, pos, id, U);
  const double rho = a.M[2 * P + id];
  const double Jp = MESH ? (double)a.G[9 * P + id] : 1.0;
  double rb[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) rb[c] = adj_inc(j.DFb[c * P + id], Bdt, j.Fb[c * P + id]);
  double gr[3][3];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int c = 0; c < 3; ++c) gr[m][c] = U[m][c];
  const double S[3] = {(double)U[0][3] + U[1][6] + U[2][7], (double)U[0][6] + U[1][4] + U[2][8],
                       (double)U[0][7] + U[1][8] + U[2][5]};
  const double e[6] = {gr[0][0], gr[1][1], gr[2][2], gr[1][0] + gr[0][1], gr[2][0] + gr[0][2],
                       gr[2][1] + gr[1][2]};
  const double tr = e[0] + e[1] + e[2];
  // velocity rates S_c/(rho J): d/drho = -rate/rho
  double g_rho = 0.0, g_lam = 0.0, g_mu = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) g_rho -= rb[c] * (S[c] / (rho * Jp)) / rho;
  // Hooke stress rates (RHS_Interior.f90:244-254)
  g_lam += (rb[3] + rb[4] + rb[5]) * tr;
  g_mu += 2.0 * (rb[3] * e[0] + rb[4] * e[1] + rb[5] * e[2]) + rb[6] * e[3] + rb[7] * e[4] +
          rb[8] * e[5];
  // memory-variable rates (w_i src_c(e) - eta)/tau_i
  if (NMECH > 0) {
    const double Qs = a.Qs[id], Qp = a.Qp[id];
#pragma unroll
    for (int i = 0; i < NMECH; ++i) {
      const double wt = a.wgt[i] * a.tau_inv[i];
      if (wt == 0.0) continue;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const long long pi = (long long)(c * NMECH + i) * P + id;
        const double re = (double)adj_inc(j.DEb[pi], Bdt, j.Eb[pi]);
        if (c < 3) {
          g_lam += re * wt * Qp * tr;
          g_mu += re * wt * (2.0 * Qs * e[c] + (2.0 * Qp - 2.0 * Qs) * tr);
        } else {
          g_mu += re * wt * Qs * e[c];
        }
      }
    }
  }
  j.g[0][id] += g_rho;
  j.g[1][id] += g_lam;
  j.g[2][id] += g_mu;
}

// true when the node lies on a face with a boundary condition (SAT penalty)
template <typename T, typename St>
__device__ __forceinline__ bool on_sat_face(const StepArgs<T, St>& a, int x, int y, int z) {
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    long long fidx, NF;
    if (a.bc[f] > 0 && on_face(a, f, x, y, z, &fidx, &NF)) return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// The node's own part of the reverse stage (no neighbour is read):
//   rb  += Bdt ub                       (reverse of u <- u + Bdt r)
//   pointwise part of K^T:
//     Z[0..8]  : curvilinear q_c = rb_v[c]/(rho J) in Z[c]; Cartesian
//                q_{d,c} = G_d rb_v[c]/rho in Z[3d+c]
//     Z[9..17] : h_{d,c} = sum_m X_m(d) gb[m][c] (Cartesian G_d gb[d][c]) with
//                gb the adjoint of the velocity gradient (Hooke + memory
//                variables)
//     etab += -rb_sigma - rb_eta/tau
//   rb   = A rb                         (reverse of r <- A r + L)
// At nodes of SAT faces rb is left unscaled for k_adj_sat, which adds the
// penalties' terms and scales it.
// ---------------------------------------------------------------------------
template <typename T, typename St, int MESH, int NMECH>
__global__ __launch_bounds__(256) void k_adj_pre(const StepArgs<T, St> a, const AdjArgs<T> j,
                                                 double A, double Bdt) {
  int x, y, z;
  const long long id = (long long)xcd_logical(blockIdx.x, gridDim.x, j.xcd) * 256 + threadIdx.x;
  if (!flat_node(a, id, x, y, z)) return;
  const long long P = a.P;
  const T lam = a.M[id], mu = a.M[P + id], rho = a.M[2 * P + id];
  const T l2m = lam + T(2) * mu;
  const double As = on_sat_face(a, x, y, z) ? 1.0 : A;
  T rb[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    const double r = adj_inc(j.DFb[c * P + id], Bdt, j.Fb[c * P + id]);
    j.DFb[c * P + id] = scale_rate(As, r);
    rb[c] = (T)r;
  }
  // adjoint of the velocity gradient gb[m][c] = d/d(dv_c/dx_m)
  T gb[3][3];
  gb[0][0] = l2m * rb[3] + lam * (rb[4] + rb[5]);
  gb[1][1] = l2m * rb[4] + lam * (rb[3] + rb[5]);
  gb[2][2] = l2m * rb[5] + lam * (rb[3] + rb[4]);
  gb[1][0] = gb[0][1] = mu * rb[6];
  gb[2][0] = gb[0][2] = mu * rb[7];
  gb[2][1] = gb[1][2] = mu * rb[8];
  if (NMECH > 0) {
    const T Qs = a.Qs[id], Qp = a.Qp[id];
    const T twomuQs = T(2) * mu * Qs, pfac = l2m * Qp - twomuQs, muQs = mu * Qs;
    T eb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, trb = T(0);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
      for (int i = 0; i < NMECH; ++i) {
        const long long pi = (long long)(c * NMECH + i) * P + id;
        const T eu = j.Eb[pi];
        const T re = adj_inc(j.DEb[pi], Bdt, eu);
        const T ti = (T)a.tau_inv[i], w = (T)a.wgt[i];
        // eta_{c,i} enters DU_sigma[c] with -1 and its own rate with -1/tau
        j.Eb[pi] = eu - rb[3 + c] - re * ti;
        j.DEb[pi] = scale_rate(A, re);
        if (c < 3) {
          eb[c] += re * ti * w * twomuQs;
          trb += re * ti * w * pfac;
        } else {
          eb[c] += re * ti * w * muQs;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) eb[c] += trb;
    gb[0][0] += eb[0];
    gb[1][1] += eb[1];
    gb[2][2] += eb[2];
    gb[1][0] += eb[3]; gb[0][1] += eb[3];
    gb[2][0] += eb[4]; gb[0][2] += eb[4];
    gb[2][1] += eb[5]; gb[1][2] += eb[5];
  }
  if (MESH) {
    const T J = a.G[9 * P + id];
    const T iq = T(1) / (rho * J);
#pragma unroll
    for (int c = 0; c < 3; ++c) j.Z[c * P + id] = rb[c] * iq;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T X0 = a.G[(0 + d) * P + id], X1 = a.G[(3 + d) * P + id], X2 = a.G[(6 + d) * P + id];
#pragma unroll
      for (int c = 0; c < 3; ++c)
        j.Z[(9 + 3 * d + c) * P + id] = X0 * gb[0][c] + X1 * gb[1][c] + X2 * gb[2][c];
    }
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T Gd = a.G[d * P + id];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        j.Z[(3 * d + c) * P + id] = Gd * rb[c] / rho;
        j.Z[(9 + 3 * d + c) * P + id] = Gd * gb[d][c];
      }
    }
  }
}

// node t of the block's boundary layer, every node once: the two x faces
// whole, the y faces without their x edges, the z faces without x and y edges
template <typename T, typename St>
__device__ __forceinline__ bool boundary_node(const StepArgs<T, St>& a, long long t, int& x, int& y,
                                              int& z) {
  const long long nx = a.nx, ny = a.ny, nz = a.nz;
  const long long nX = ny * nz, nY = (nx - 2) * nz, nZ = (nx - 2) * (ny - 2);
  if (t < 2 * nX) {
    const long long k = t >= nX ? t - nX : t;
    x = t >= nX ? (int)nx - 1 : 0;
    y = (int)(k % ny);
    z = (int)(k / ny);
    return true;
  }
  t -= 2 * nX;
  if (t < 2 * nY) {
    const long long k = t >= nY ? t - nY : t;
    y = t >= nY ? (int)ny - 1 : 0;
    x = 1 + (int)(k % (nx - 2));
    z = (int)(k / (nx - 2));
    return true;
  }
  t -= 2 * nY;
  if (t < 2 * nZ) {
    const long long k = t >= nZ ? t - nZ : t;
    z = t >= nZ ? (int)nz - 1 : 0;
    x = 1 + (int)(k % (nx - 2));
    y = 1 + (int)(k / (nx - 2));
    return true;
  }
  return false;
}

// ---------------------------------------------------------------------------
// SAT penalties of the nodes on boundary-condition faces (after k_adj_pre,
// which left rb = DFb + Bdt ub unscaled there):
//   g  += (dL/dm)^T rb of the penalties at the stage input state (dual numbers)
//   Fb += S^T rb (the penalty is linear in the node's own state, S u; its
//         transpose is assembled column by column, S e_k)
//   rb  = A rb
// One thread per boundary node, so a node on several faces sums them all.
// ---------------------------------------------------------------------------
template <typename T, typename St>
__global__ __launch_bounds__(256) void k_adj_sat(const StepArgs<T, St> a, const AdjArgs<T> j,
                                                 double A) {
  int x, y, z;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (!boundary_node(a, t, x, y, z)) return;
  if (!on_sat_face(a, x, y, z)) return;
  const long long P = a.P;
  const long long id = x + (long long)a.nx * (y + (long long)a.ny * z);
  const double lam = a.M[id], mu = a.M[P + id], rho = a.M[2 * P + id];
  const T lamT = (T)lam, muT = (T)mu, rhoT = (T)rho;
  double rb[9];
  T rbT[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    rb[c] = j.DFb[c * P + id];
    rbT[c] = (T)rb[c];
  }
  T sb[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) sb[k] = T(0);
  double g_rho = 0.0, g_lam = 0.0, g_mu = 0.0;
  for (int f = 0; f < 6; ++f) {
    if (a.bc[f] <= 0) continue;
    long long fidx, NF;
    if (!on_face(a, f, x, y, z, &fidx, &NF)) continue;
    for (int k = 0; k < 9; ++k) {
      T u[9], out[9];
#pragma unroll
      for (int c = 0; c < 9; ++c) u[c] = c == k ? T(1) : T(0);
      sat_rate<T, T>(a, f, fidx, NF, u, lamT, muT, rhoT, out);
      T s = T(0);
#pragma unroll
      for (int c = 0; c < 9; ++c) s += out[c] * rbT[c];
      sb[k] += s;
    }
    D3 u[9], out[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) u[c] = D3((double)(T)a.Fin[c * P + id]);
    D3 dl(lam), dm(mu), dr(rho);
    dr.d[0] = 1.0;
    dl.d[1] = 1.0;
    dm.d[2] = 1.0;
    sat_rate<D3, T>(a, f, fidx, NF, u, dl, dm, dr, out);
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      g_rho += rb[c] * out[c].d[0];
      g_lam += rb[c] * out[c].d[1];
      g_mu += rb[c] * out[c].d[2];
    }
  }
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    j.Fb[k * P + id] += (double)sb[k];
    j.DFb[k * P + id] = scale_rate(A, rb[k]);
  }
  j.g[0][id] += g_rho;
  j.g[1][id] += g_lam;
  j.g[2][id] += g_mu;
}

// ---------------------------------------------------------------------------
// transposed derivatives: Fb += (D^T applied to Z)
//   curvilinear: sigma_b(m,c) += J X_m(d) / h_d (D_d^T Z[c]);  vb_c += 1/h_d D_d^T Z[9+3d+c]
//   Cartesian:   sigma_b(d,c) += 1/h_d D_d^T Z[3d+c];         vb_c as above
// ---------------------------------------------------------------------------
template <typename T, typename St, int MESH>
__global__ __launch_bounds__(256) void k_adj_stencil(const StepArgs<T, St> a, const AdjArgs<T> j) {
  int x, y, z;
  const long long id = (long long)xcd_logical(blockIdx.x, gridDim.x, j.xcd) * 256 + threadIdx.x;
  if (!flat_node(a, id, x, y, z)) return;
  const long long P = a.P;
  const int pos[3] = {x, y, z};
  const int n[3] = {a.nx, a.ny, a.nz};
  const long long st[3] = {1, a.nx, (long long)a.nx * a.ny};
  double vb[3] = {0.0, 0.0, 0.0};
  double sb[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T zq[3] = {T(0), T(0), T(0)}, zh[3] = {T(0), T(0), T(0)};
    for_taps_T(pos[d], n[d], [&](int off, double cd) {
      const long long p = id + (long long)off * st[d];
      const T ck = (T)cd;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zq[c] += ck * j.Z[((MESH ? 0 : 3 * d) + c) * P + p];
        zh[c] += ck * j.Z[(9 + 3 * d + c) * P + p];
      }
    });
    const double ih = a.ih[d];
#pragma unroll
    for (int c = 0; c < 3; ++c) vb[c] += ih * (double)zh[c];
    if (MESH) {
      const double J = a.G[9 * P + id];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double JX = J * (double)a.G[(3 * m + d) * P + id] * ih;
#pragma unroll
        for (int c = 0; c < 3; ++c) sb[m][c] += JX * (double)zq[c];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) sb[d][c] += ih * (double)zq[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) j.Fb[c * P + id] += vb[c];
  j.Fb[3 * P + id] += sb[0][0];
  j.Fb[4 * P + id] += sb[1][1];
  j.Fb[5 * P + id] += sb[2][2];
  j.Fb[6 * P + id] += sb[0][1] + sb[1][0];
  j.Fb[7 * P + id] += sb[0][2] + sb[2][0];
  j.Fb[8 * P + id] += sb[1][2] + sb[2][1];
}

// Fb at receiver nodes += g[r][c]
__global__ void k_adj_inject(double* Fb, long long P, const long long* idx, int n, const double* g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 9 * n) return;
  const int r = t / 9, c = t - 9 * (t / 9);
  // receivers at one node add up: one thread per (receiver, component) pair
  // may share a node with another receiver, so accumulate atomically
  atomicAdd(&Fb[c * P + idx[r]], g[t]);
}

// ---------------------------------------------------------------------------
template <typename T, typename St>
hipError_t launch_adjoint_stage(const StepArgs<T, St>& a, const AdjArgs<T>& j, int mesh, double A,
                                double Bdt, hipStream_t s) {
  const long long nodes = (long long)a.nx * a.ny * a.nz;
  const dim3 g((unsigned)((nodes + 255) / 256)), b(256);
  // boundary layer: the x faces whole, y faces without x edges, z faces
  // without x and y edges (block sizes >= 2 per axis)
  const long long nb = 2LL * a.ny * a.nz + 2LL * (a.nx - 2) * a.nz + 2LL * (a.nx - 2) * (a.ny - 2);
  const dim3 gb((unsigned)((nb + 255) / 256));
  hipError_t e = hipSuccess;
#define WQL_CASE(MS, NM)                                                       \
  if (mesh == MS && a.nmech == NM) {                                           \
    hipLaunchKernelGGL((k_adj_grad<T, St, MS, NM>), g, b, 0, s, a, j, Bdt);    \
    hipLaunchKernelGGL((k_adj_pre<T, St, MS, NM>), g, b, 0, s, a, j, A, Bdt);  \
    hipLaunchKernelGGL((k_adj_sat<T, St>), gb, b, 0, s, a, j, A);              \
    hipLaunchKernelGGL((k_adj_stencil<T, St, MS>), g, b, 0, s, a, j);          \
    e = hipGetLastError();                                                     \
  } else
  WQL_CASE(0, 0) WQL_CASE(0, 4) WQL_CASE(0, 8) WQL_CASE(1, 0) WQL_CASE(1, 4) WQL_CASE(1, 8)
  return hipErrorInvalidValue;
#undef WQL_CASE
  return e;
}
template hipError_t launch_adjoint_stage<double, double>(const StepArgs<double, double>&,
                                                         const AdjArgs<double>&, int, double,
                                                         double, hipStream_t);
template hipError_t launch_adjoint_stage<float, double>(const StepArgs<float, double>&,
                                                        const AdjArgs<float>&, int, double,
                                                        double, hipStream_t);
template hipError_t launch_adjoint_stage<float, float>(const StepArgs<float, float>&,
                                                       const AdjArgs<float>&, int, double,
                                                       double, hipStream_t);

hipError_t launch_adjoint_inject(double* Fb, long long P, const long long* idx, int n,
                                 const double* g, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_adj_inject, dim3((9 * n + 255) / 256), dim3(256), 0, s, Fb, P, idx, n, g);
  return hipGetLastError();
}

}  // namespace wql
