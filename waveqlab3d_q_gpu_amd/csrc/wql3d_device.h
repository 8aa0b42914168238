// Device helpers shared by the forward kernels (wql3d_kernels.hip) and the
// discrete-adjoint kernels (wql3d_adjoint.hip): the SBP-6 operator tables and
// tap iterators, plane addressing, the derivative evaluators, PML slab
// geometry, the SAT boundary data and the moment-tensor source.
//
// The SBP operator is table driven: rows 0..5 and n-6..n-1 of every axis use
// the 6x9 closure table (reference JU_xJU_yJU_z6.f90:74-299), other rows the
// 6-tap central stencil (JU_xJU_yJU_z6.f90:53-59).
#pragma once

#include <type_traits>

#include "wql3d_internal.h"

namespace wql {

// SBP-6 left closure rows (columns 0..8), reference JU_xJU_yJU_z6.f90:74-182.
// The literal 0.013435342414630 (row 3, column 6) is the double value, as in
// the reference's cpu-release build with -fdefault-real-8.
static __constant__ double kSbpL[6][9] = {
    {-1.582533518939116, 2.033378678700676, -0.141512858744873, -0.450398306578272,
     0.104488069284042, 0.036577936277544, 0.0, 0.0, 0.0},
    {-0.462059195631158, 0.0, 0.287258622978251, 0.258816087376832, -0.069112065532624,
     -0.014903449191300, 0.0, 0.0, 0.0},
    {0.071247104721830, -0.636451095137907, 0.0, 0.606235523609147, -0.022902190275815,
     -0.018129342917256, 0.0, 0.0, 0.0},
    {0.114713313798970, -0.290087484386815, -0.306681191361148, 0.0, 0.520262285050482,
     -0.051642265516119, 0.013435342414630, 0.0, 0.0},
    {-0.036210680656541, 0.105400944933782, 0.015764336127392, -0.707905442575989, 0.0,
     0.769199413962647, -0.164529643265203, 0.018281071473911, 0.0},
    {-0.011398193015050, 0.020437334208704, 0.011220896474665, 0.063183694641876,
     -0.691649024426814, 0.0, 0.739709139060752, -0.147941827812150, 0.016437980868017},
};

// singular_source order-4 stencils (moment_tensor.f90:980-1121), alpha 0,25,50,100
static __constant__ double kSrc4[4][9] = {
    {-0.01171875, -0.03125, 0.046875, 0.28125, 0.4296875, 0.28125, 0.046875, -0.03125,
     -0.01171875},
    {-0.00246875, -0.02275, -0.014125, 0.16575, 0.4171875, 0.38675, 0.117875, -0.02975,
     -0.01846875},
    {-0.00728125, -0.028, 0.014875, 0.224, 0.4265625, 0.336, 0.081375, -0.032, -0.01553125},
    {-0.0095625, -0.02990625, 0.03053125, 0.25284375, 0.42890625, 0.30903125, 0.06384375,
     -0.03196875, -0.01371875},
};

// Apply f(offset, coefficient) for every tap of SBP D1 row i of an n-point axis.
template <typename Fn>
__device__ __forceinline__ void for_taps(int i, int n, Fn&& f) {
  if (i >= 6 && i < n - 6) {
    f(-3, -1.0 / 60.0);
    f(-2, 0.15);
    f(-1, -0.75);
    f(1, 0.75);
    f(2, -0.15);
    f(3, 1.0 / 60.0);
  } else if (i < 6) {
#pragma unroll
    for (int k = 0; k < 9; ++k) f(k - i, kSbpL[i][k]);
  } else {
    const int r = n - 1 - i;
#pragma unroll
    for (int k = 0; k < 9; ++k) f(n - 9 + k - i, -kSbpL[r][8 - k]);
  }
}

// The central taps only (rows 6..n-7), with compile-time offsets.
template <typename Fn>
__device__ __forceinline__ void for_taps_interior(Fn&& f) {
  f(-3, -1.0 / 60.0);
  f(-2, 0.15);
  f(-1, -0.75);
  f(1, 0.75);
  f(2, -0.15);
  f(3, 1.0 / 60.0);
}

__device__ __forceinline__ double scale_rate(double A, double x) { return A == 0.0 ? 0.0 : A * x; }
__device__ __forceinline__ float scale_rate(double A, float x) {
  return A == 0.0 ? 0.0f : (float)A * x;
}

// ---------------------------------------------------------------------------
// U[m][c] = JJU_x6 (JU_xJU_yJU_z6.f90:12-809) evaluated with metric array m
// (0 metricx, 1 metricy, 2 metricz): c<3 -> sum_d X_m(pt,d)/h_d D_d(v_c);
// c>=3 -> sum_d 1/h_d D_d(J X_m(.,d) sigma_c).
// ---------------------------------------------------------------------------
// Node offsets are 32-bit (a plane holds < 2^29 nodes, so byte offsets fit in
// 32 bits) and every plane is addressed from its own wavefront-uniform base
// pointer: the load then takes a scalar base plus one 32-bit vector offset,
// shared by all plane loads of a tap.
template <typename V>
__device__ __forceinline__ V ldp(const V* base, unsigned idx) {
  return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(V));
}
template <typename V>
__device__ __forceinline__ V* atp(V* base, unsigned idx) {
  using Byte = typename std::conditional<std::is_const<V>::value, const char, char>::type;
  return reinterpret_cast<V*>(reinterpret_cast<Byte*>(base) + idx * (unsigned)sizeof(V));
}
// the planes the stencils difference: F_in, or its float copy (FP32-mixed)
template <typename T, typename S>
__device__ __forceinline__ const OpType<T, S>* op_planes(const StepArgs<T, S>& a) {
  if constexpr (kOpCopy<T, S>) return a.Fo_in;
  else return a.Fin;
}

template <typename T, typename S>
__device__ __forceinline__ void derivs_curv(const StepArgs<T, S>& a, const int pos[3],
                                            long long id64, T U[3][9]) {
  const long long P = a.P;
  const unsigned id = (unsigned)id64;
  const unsigned st[3] = {1u, (unsigned)a.nx, (unsigned)a.nx * (unsigned)a.ny};
  const int n[3] = {a.nx, a.ny, a.nz};
  const OpType<T, S>* Fc[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) Fc[c] = op_planes(a) + c * P;
  const T* Gk[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) Gk[k] = a.G + k * P;
  T dv[3][3];
  T Sf[3][6];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int c = 0; c < 6; ++c) Sf[m][c] = T(0);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const T ih = (T)a.ih[d];
    T v0 = T(0), v1 = T(0), v2 = T(0);
    for_taps(pos[d], n[d], [&](int off, double cd) {
      const unsigned p = id + (unsigned)off * st[d];
      const T ck = ih * (T)cd;
      v0 += ck * (T)ldp(Fc[0], p);
      v1 += ck * (T)ldp(Fc[1], p);
      v2 += ck * (T)ldp(Fc[2], p);
      const T j = ldp(Gk[9], p);
      const T jm0 = j * ldp(Gk[0 + d], p);
      const T jm1 = j * ldp(Gk[3 + d], p);
      const T jm2 = j * ldp(Gk[6 + d], p);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        const T t = ck * (T)ldp(Fc[3 + c], p);
        Sf[0][c] += jm0 * t;
        Sf[1][c] += jm1 * t;
        Sf[2][c] += jm2 * t;
      }
    });
    dv[d][0] = v0;
    dv[d][1] = v1;
    dv[d][2] = v2;
  }
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const T X0 = ldp(Gk[3 * m + 0], id), X1 = ldp(Gk[3 * m + 1], id),
            X2 = ldp(Gk[3 * m + 2], id);
#pragma unroll
    for (int c = 0; c < 3; ++c) U[m][c] = X0 * dv[0][c] + X1 * dv[1][c] + X2 * dv[2][c];
#pragma unroll
    for (int c = 0; c < 6; ++c) U[m][3 + c] = Sf[m][c];
  }
}

// Cartesian mesh: diagonal metrics only, J-free form of the interior branch
// (JU_xJU_yJU_z6.f90:16426-16470): U[d][c] = X_d(pt) / h_d * D_d(F_c).
template <typename T, typename S>
__device__ __forceinline__ void derivs_cart(const StepArgs<T, S>& a, const int pos[3],
                                            long long id64, T U[3][9]) {
  const long long P = a.P;
  const unsigned id = (unsigned)id64;
  const unsigned st[3] = {1u, (unsigned)a.nx, (unsigned)a.nx * (unsigned)a.ny};
  const int n[3] = {a.nx, a.ny, a.nz};
  const OpType<T, S>* Fc[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) Fc[c] = op_planes(a) + c * P;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    T acc[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) acc[c] = T(0);
    for_taps(pos[d], n[d], [&](int off, double cd) {
      const unsigned p = id + (unsigned)off * st[d];
      const T ck = (T)cd;
#pragma unroll
      for (int c = 0; c < 9; ++c) acc[c] += ck * (T)ldp(Fc[c], p);
    });
    const T s = a.G[d * P + id] * (T)a.ih[d];
#pragma unroll
    for (int c = 0; c < 9; ++c) U[d][c] = s * acc[c];
  }
}

// ---------------------------------------------------------------------------
// PML geometry (pml.f90:11-168; RHS_Interior.f90:1846-2451)
// ---------------------------------------------------------------------------
template <typename T, typename S>
__device__ __forceinline__ bool pml_slab(const StepArgs<T, S>& a, int f, int x, int y, int z,
                                         long long* sidx, double* d) {
  const int np = a.npml[f];
  const double d0 = 3.0 * 6.907755278982137;  // 6/(2*1)*log(1/Rf), Rf = 1e-3
  double r;
  switch (f) {
    case 0:
      if (x >= np) return false;
      *sidx = x + (long long)np * (y + (long long)a.ny * z);
      r = (double)(x - np);              // x1 - (1+np), x1 = x+1
      break;
    case 1:
      if (x < a.nx - np) return false;
      *sidx = (x - (a.nx - np)) + (long long)np * (y + (long long)a.ny * z);
      r = (double)(x + 1 - (a.nx - np));
      break;
    case 2:
      if (y >= np) return false;
      *sidx = x + (long long)a.nx * (y + (long long)np * z);
      r = (double)(y - np);
      break;
    case 3:
      if (y < a.ny - np) return false;
      *sidx = x + (long long)a.nx * ((y - (a.ny - np)) + (long long)np * z);
      r = (double)(y + 1 - (a.ny - np));
      break;
    case 4:
      if (z >= np) return false;
      *sidx = x + (long long)a.nx * (y + (long long)a.ny * z);
      r = (double)(z + 1 - np);          // z-low profile |z - npml| (RHS_Interior.f90:1919)
      break;
    default:
      if (z < a.nz - np) return false;
      *sidx = x + (long long)a.nx * (y + (long long)a.ny * (z - (a.nz - np)));
      r = (double)(z + 1 - (a.nz - np));
      break;
  }
  *d = d0 * fabs(r / (double)np);
  return true;
}

template <typename T, typename S>
__device__ __forceinline__ long long slab_plane(const StepArgs<T, S>& a, int f) {
  const long long np = a.npml[f];
  return f < 2 ? np * a.ny * a.nz : (f < 4 ? (long long)a.nx * np * a.nz : (long long)a.nx * a.ny * np);
}

// ---------------------------------------------------------------------------
// SAT boundary data ubc at a face node: BCm/BCp (BoundaryConditions.f90:783-924)
// ---------------------------------------------------------------------------
// Frame of face f at face index fidx: l[3], m[3], n[3] and |grad xi|.
template <typename T, typename S>
__device__ __forceinline__ void face_frame(const StepArgs<T, S>& a, int f, long long fidx, long long NF,
                                           T fr[10]) {
  const T* fd = a.face[f];
#pragma unroll
  for (int k = 0; k < 10; ++k) fr[k] = fd[k * NF + fidx];
}

// ubc of one face node from its frame.  S is the arithmetic type (the
// storage type T, or a dual number carrying material derivatives), R the
// frame type.
template <typename S, typename R>
__device__ __forceinline__ void sat_ubc_frame(const R fr[10], int code, int f, const S u[9], S lam,
                                              S mu, S rho, S ubc[9]) {
  const R l0 = fr[0], l1 = fr[1], l2 = fr[2];
  const R m0 = fr[3], m1 = fr[4], m2 = fr[5];
  const R n0 = fr[6], n1 = fr[7], n2 = fr[8];
  const R norm = fr[9];
  const S r = code == 1 ? S(0) : (code == 2 ? S(1) : S(-1));
  const S cp = sqrt((S(2) * mu + lam) / rho), cs = sqrt(mu / rho);
  const S Zp = rho * cp, Zs = rho * cs;
  const S T0 = n0 * u[3] + n1 * u[6] + n2 * u[7];
  const S T1 = n0 * u[6] + n1 * u[4] + n2 * u[8];
  const S T2 = n0 * u[7] + n1 * u[8] + n2 * u[5];
  const S vl = l0 * u[0] + l1 * u[1] + l2 * u[2];
  const S vm = m0 * u[0] + m1 * u[1] + m2 * u[2];
  const S vn = n0 * u[0] + n1 * u[1] + n2 * u[2];
  const S Tl = l0 * T0 + l1 * T1 + l2 * T2;
  const S Tm = m0 * T0 + m1 * T1 + m2 * T2;
  const S Tn = n0 * T0 + n1 * T1 + n2 * T2;
  const S sg = (f & 1) ? S(1) : S(-1);   // BCp on high faces, BCm on low faces
  const S a1 = S(0.5) * (S(1) - r), a2 = S(0.5) * (S(1) + r) * sg;
  const S pl = a1 * Zs * vl + a2 * Tl;
  const S pm = a1 * Zs * vm + a2 * Tm;
  const S pn = a1 * Zp * vn + a2 * Tn;
  const S ql = a1 * vl + a2 / Zs * Tl;
  const S qm = a1 * vm + a2 / Zs * Tm;
  const S qn = a1 * vn + a2 / Zp * Tn;
  const S Fx = l0 * pl + m0 * pm + n0 * pn;
  const S Fy = l1 * pl + m1 * pm + n1 * pn;
  const S Fz = l2 * pl + m2 * pm + n2 * pn;
  const S Gx = l0 * ql + m0 * qm + n0 * qn;
  const S Gy = l1 * ql + m1 * qm + n1 * qn;
  const S Gz = l2 * ql + m2 * qm + n2 * qn;
  const S sn = sg * norm;
  ubc[0] = norm * Fx;
  ubc[1] = norm * Fy;
  ubc[2] = norm * Fz;
  ubc[3] = sn * n0 * Gx;
  ubc[4] = sn * n1 * Gy;
  ubc[5] = sn * n2 * Gz;
  ubc[6] = sn * (n1 * Gx + n0 * Gy);
  ubc[7] = sn * (n2 * Gx + n0 * Gz);
  ubc[8] = sn * (n2 * Gy + n1 * Gz);
}

template <typename T, typename S>
__device__ __forceinline__ void sat_ubc(const StepArgs<T, S>& a, int f, long long fidx, long long NF,
                                        const T u[9], T lam, T mu, T rho, T ubc[9]) {
  T fr[10];
  face_frame(a, f, fidx, NF, fr);
  sat_ubc_frame<T, T>(fr, a.bc[f], f, u, lam, mu, rho, ubc);
}

// structure(ubc) (RHS_Interior.f90:2267-2287): the rate vector of a SAT penalty
template <typename S>
__device__ __forceinline__ void sat_structure(const S ubc[9], S lam, S mu, S rho, S Ux[9]) {
  Ux[0] = S(1) / rho * ubc[0];
  Ux[1] = S(1) / rho * ubc[1];
  Ux[2] = S(1) / rho * ubc[2];
  Ux[3] = (S(2) * mu + lam) * ubc[3] + lam * (ubc[4] + ubc[5]);
  Ux[4] = (S(2) * mu + lam) * ubc[4] + lam * (ubc[3] + ubc[5]);
  Ux[5] = (S(2) * mu + lam) * ubc[5] + lam * (ubc[3] + ubc[4]);
  Ux[6] = mu * ubc[6];
  Ux[7] = mu * ubc[7];
  Ux[8] = mu * ubc[8];
}

// The alpha/theta form of the boundary penalty (commented out at
// BoundaryConditions.f90:833-841 and 905-913; alpha, theta per code as set in
// BC_Lx..BC_Rz, BoundaryConditions.f90:129-160): the rate vector itself,
// velocity alpha*norm/rho*F, stresses theta*norm*(...), no structure().
template <typename S, typename R>
__device__ __forceinline__ void sat_alpha_theta_frame(const R fr[10], int code, int f,
                                                      const S u[9], S lam, S mu, S rho,
                                                      S Ux[9]) {
  const R l0 = fr[0], l1 = fr[1], l2 = fr[2];
  const R m0 = fr[3], m1 = fr[4], m2 = fr[5];
  const R n0 = fr[6], n1 = fr[7], n2 = fr[8];
  const R norm = fr[9];
  const S r = code == 1 ? S(0) : (code == 2 ? S(1) : S(-1));
  const S alpha = code == 1 ? S(2) : (code == 2 ? S(1) : S(0));
  const S theta = code == 3 ? S(-1) : S(0);
  const S cp = sqrt((S(2) * mu + lam) / rho), cs = sqrt(mu / rho);
  const S Zp = rho * cp, Zs = rho * cs;
  const S T0 = n0 * u[3] + n1 * u[6] + n2 * u[7];
  const S T1 = n0 * u[6] + n1 * u[4] + n2 * u[8];
  const S T2 = n0 * u[7] + n1 * u[8] + n2 * u[5];
  const S vl = l0 * u[0] + l1 * u[1] + l2 * u[2];
  const S vm = m0 * u[0] + m1 * u[1] + m2 * u[2];
  const S vn = n0 * u[0] + n1 * u[1] + n2 * u[2];
  const S Tl = l0 * T0 + l1 * T1 + l2 * T2;
  const S Tm = m0 * T0 + m1 * T1 + m2 * T2;
  const S Tn = n0 * T0 + n1 * T1 + n2 * T2;
  const S sg = (f & 1) ? S(1) : S(-1);
  const S a1 = S(0.5) * (S(1) - r), a2 = S(0.5) * (S(1) + r) * sg;
  const S pl = a1 * Zs * vl + a2 * Tl;
  const S pm = a1 * Zs * vm + a2 * Tm;
  const S pn = a1 * Zp * vn + a2 * Tn;
  const S ql = a1 * vl + a2 / Zs * Tl;
  const S qm = a1 * vm + a2 / Zs * Tm;
  const S qn = a1 * vn + a2 / Zp * Tn;
  const S Fx = l0 * pl + m0 * pm + n0 * pn;
  const S Fy = l1 * pl + m1 * pm + n1 * pn;
  const S Fz = l2 * pl + m2 * pm + n2 * pn;
  const S Gx = l0 * ql + m0 * qm + n0 * qn;
  const S Gy = l1 * ql + m1 * qm + n1 * qn;
  const S Gz = l2 * ql + m2 * qm + n2 * qn;
  const S l2m = lam + S(2) * mu;
  Ux[0] = alpha * norm / rho * Fx;
  Ux[1] = alpha * norm / rho * Fy;
  Ux[2] = alpha * norm / rho * Fz;
  Ux[3] = theta * norm * (l2m * n0 * Gx + lam * (n1 * Gy + n2 * Gz));
  Ux[4] = theta * norm * (l2m * n1 * Gy + lam * (n0 * Gx + n2 * Gz));
  Ux[5] = theta * norm * (l2m * n2 * Gz + lam * (n0 * Gx + n1 * Gy));
  Ux[6] = theta * mu * norm * (n1 * Gx + n0 * Gy);
  Ux[7] = theta * mu * norm * (n2 * Gx + n0 * Gz);
  Ux[8] = theta * mu * norm * (n2 * Gy + n1 * Gz);
}

// rate vector of the boundary penalty at a face node in the block's SAT form
// (DF -= pen * Ux); ubc (the PML DQ penalty) is filled for the BCm/BCp form
template <typename T, typename S>
__device__ __forceinline__ void sat_rate(const StepArgs<T, S>& a, int f, long long fidx, long long NF,
                                         const T u[9], T lam, T mu, T rho, T ubc[9], T Ux[9]) {
  T fr[10];
  face_frame(a, f, fidx, NF, fr);
  if (a.sat_form == WQL_SAT_ALPHA_THETA) {
    sat_alpha_theta_frame<T, T>(fr, a.bc[f], f, u, lam, mu, rho, Ux);
  } else {
    sat_ubc_frame<T, T>(fr, a.bc[f], f, u, lam, mu, rho, ubc);
    sat_structure<T>(ubc, lam, mu, rho, Ux);
  }
}

// face index of node (x,y,z) on face f; returns false if not on the face
template <typename T, typename S>
__device__ __forceinline__ bool on_face(const StepArgs<T, S>& a, int f, int x, int y, int z,
                                        long long* fidx, long long* NF) {
  switch (f) {
    case 0: *fidx = y + (long long)a.ny * z; *NF = (long long)a.ny * a.nz; return x == 0;
    case 1: *fidx = y + (long long)a.ny * z; *NF = (long long)a.ny * a.nz; return x == a.nx - 1;
    case 2: *fidx = x + (long long)a.nx * z; *NF = (long long)a.nx * a.nz; return y == 0;
    case 3: *fidx = x + (long long)a.nx * z; *NF = (long long)a.nx * a.nz; return y == a.ny - 1;
    case 4: *fidx = x + (long long)a.nx * y; *NF = (long long)a.nx * a.ny; return z == 0;
    default: *fidx = x + (long long)a.nx * y; *NF = (long long)a.nx * a.ny; return z == a.nz - 1;
  }
}

// singular_source (moment_tensor.f90:938-1124)
__device__ __forceinline__ double singular_source(int i, int i0, double h, double aa, int order,
                                                  int alpha) {
  if (order == 2) {
    const double beta = (1.0 - 2.0 * aa) / 4.0 + 0.25;
    const double gam = (2.0 * aa - 1.0) / 16.0;
    const double gap = (2.0 * aa + 1.0) / 16.0;
    if (i == i0 - 2) return (-gam) / h;
    if (i == i0 - 1) return (beta + 3.0 * gam - gap) / h;
    if (i == i0) return (1.0 - aa - 2.0 * beta - 3.0 * gam + 3.0 * gap) / h;
    if (i == i0 + 1) return (aa + beta + gam - 3.0 * gap) / h;
    if (i == i0 + 2) return gap / h;
    return 0.0;
  }
  const int k = i - i0 + 4;
  if (k < 0 || k > 8) return 0.0;
  const int row = alpha == 25 ? 1 : (alpha == 50 ? 2 : (alpha == 100 ? 3 : 0));
  return kSrc4[row][k] / h;
}

// source contribution (to subtract from DF(4..9)) at node (x,y,z)
// support half-width of source Sr in nodes (the box source_at visits)
__host__ __device__ __forceinline__ int source_radius(const DevSource& Sr) {
  return Sr.order == 0 ? kMollifyRadius : (Sr.order == 2 ? 2 : 4);
}

// set_moment_tensor_smooth (moment_tensor.f90:1255-1330) at node (x, y, z):
// a Gaussian of standard deviation 2h per axis, h the coordinate's backward
// difference along its own index (forward at the first node), with the
// reference's truncated pi
template <typename T, typename S>
__device__ __forceinline__ double mollified_delta(const StepArgs<T, S>& a, int x, int y, int z,
                                                  long long id, const DevSource& Sr) {
  const double pi = 3.14159265359;
  const long long P = a.P, sy = a.nx, sz = (long long)a.nx * a.ny;
  const double* X = a.X;
  const double hx = x == 0 ? X[id + 1] - X[id] : X[id] - X[id - 1];
  const double hy = y == 0 ? X[P + id + sy] - X[P + id] : X[P + id] - X[P + id - sy];
  const double hz = z == 0 ? X[2 * P + id + sz] - X[2 * P + id] : X[2 * P + id] - X[2 * P + id - sz];
  const double sgx = 2.0 * hx, sgy = 2.0 * hy, sgz = 2.0 * hz;
  const double ux = (X[id] - Sr.loc[0]) / sgx;
  const double uy = (X[P + id] - Sr.loc[1]) / sgy;
  const double uz = (X[2 * P + id] - Sr.loc[2]) / sgz;
  const double dx = exp(-0.5 * (ux * ux)) / (sgx * sqrt(2.0 * pi));
  const double dy = exp(-0.5 * (uy * uy)) / (sgy * sqrt(2.0 * pi));
  const double dz = exp(-0.5 * (uz * uz)) / (sgz * sqrt(2.0 * pi));
  return dx * dy * dz;
}

template <typename T, int MESH, typename S>
__device__ __forceinline__ bool source_at(const StepArgs<T, S>& a, int x, int y, int z, long long id,
                                          double out[6]) {
  bool any = false;
  for (int s = 0; s < a.nsrc; ++s) {
    const DevSource& Sr = a.src[s];
    const int r = source_radius(Sr);
    if (x < Sr.near[0] - r || x > Sr.near[0] + r || y < Sr.near[1] - r || y > Sr.near[1] + r ||
        z < Sr.near[2] - r || z > Sr.near[2] + r)
      continue;
    double delta;
    if (Sr.order == 0) {
      delta = mollified_delta(a, x, y, z, id, Sr);
    } else {
      const long long P = a.P;
      const double qx = (double)a.G[0 * P + id];
      const double ry = (double)a.G[(MESH ? 4 : 1) * P + id];
      const double sz = (double)a.G[(MESH ? 8 : 2) * P + id];
      const double hx = a.h[0] / qx, hy = a.h[1] / ry, hz = a.h[2] / sz;
      const double dx = singular_source(x, Sr.near[0], hx, Sr.off[0] / hx, Sr.order, Sr.alpha);
      const double dy = singular_source(y, Sr.near[1], hy, Sr.off[1] / hy, Sr.order, Sr.alpha);
      const double dz = singular_source(z, Sr.near[2], hz, Sr.off[2] / hz, Sr.order, Sr.alpha);
      delta = dx * dy * dz;
    }
    if (!any) {
#pragma unroll
      for (int c = 0; c < 6; ++c) out[c] = 0.0;
      any = true;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) out[c] += Sr.am[c] * delta;
  }
  return any;
}

// Streams read and written once per stage (DF, memory variables, PML
// auxiliaries, F_out, material/Q at the node) may bypass cache residency with
// non-temporal accesses so the stencil operands (F_in, geometry) keep it.
template <bool NT, typename V>
__device__ __forceinline__ V ld_stream(const V* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void st_stream(V* p, V v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Workgroup -> logical index.  Workgroups are issued round-robin over the 8
// XCDs (workgroup h to XCD h % 8).  xcd_remap = 1 permutes the index so each
// XCD receives a contiguous range of logical indices; xcd_remap = G >= 2 so
// each XCD receives groups of G consecutive logical indices (XCD j the groups
// j, j + 8, j + 16, ...; the last B mod 8G workgroups keep their index), so
// that neighbouring work shares one L2 while the XCDs stay interleaved over
// the array.  Both are bijections on [0, B).
__device__ __forceinline__ unsigned xcd_logical(unsigned h, unsigned B, int remap) {
  if (remap == 1) {
    const unsigned q = B >> 3, r = B & 7u, j = h & 7u, k = h >> 3;
    return j * q + (j < r ? j : r) + k;
  }
  if (remap >= 2) {
    const unsigned G = (unsigned)remap, B8 = B / (8u * G) * (8u * G);
    if (h < B8) {
      const unsigned j = h & 7u, k = h >> 3;
      return ((k / G) * 8u + j) * G + k % G;
    }
  }
  return h;
}

// Old rate register of the low-storage RK scheme, to be scaled by A.  The
// first stage has A = 0 and scale_rate maps any old rate to 0 there, so the
// register is not read at all (a uniform branch on a kernel argument): that
// stage moves no old-rate bytes.
template <bool NT, typename V>
__device__ __forceinline__ V ld_rate(double A, const V* p) {
  if (A == 0.0) return V(0);
  return ld_stream<NT>(p);
}

// DF = df and the RK state update F_out = F_in + Bdt*DF of component c at
// node idu: FP64 state, or FP32 state with a Kahan-compensated sum whose
// compensation planes Fc are updated in place.  DF is not stored when the
// stage's rates are dead (StepArgs::dead_rates).
template <bool NT, typename T, typename S>
__device__ __forceinline__ void rk_finish(const StepArgs<T, S>& a, int c, unsigned idu, double df) {
  const long long P = a.P;
  if constexpr (std::is_same<S, double>::value) {
    if (!a.dead_rates) st_stream<NT>(atp(a.DF + c * P, idu), df);
    const double fo = ldp(a.Fin + c * P, idu) + a.Bdt * df;
    st_stream<NT>(atp(a.Fout + c * P, idu), fo);
    if constexpr (kOpCopy<T, S>) st_stream<NT>(atp(a.Fo_out + c * P, idu), (T)fo);
  } else {
    const float d = (float)df;
    if (!a.dead_rates) st_stream<NT>(atp(a.DF + c * P, idu), d);
    const float f = ldp(a.Fin + c * P, idu);
    float* cp = atp(a.Fc + c * P, idu);
    const float y = (float)a.Bdt * d - ld_stream<NT>(cp);
    const float t = f + y;
    st_stream<NT>(cp, (t - f) - y);
    st_stream<NT>(atp(a.Fout + c * P, idu), t);
  }
}

// The same update from values loaded ahead of time: the state fin, the old
// rate dfo and (FP32 state) the compensation comp of component c.
template <bool NT, typename T, typename S>
__device__ __forceinline__ void rk_finish_pre(const StepArgs<T, S>& a, int c, unsigned idu,
                                              double df, S fin, S comp) {
  const long long P = a.P;
  if constexpr (std::is_same<S, double>::value) {
    (void)comp;
    if (!a.dead_rates) st_stream<NT>(atp(a.DF + c * P, idu), df);
    const double fo = fin + a.Bdt * df;
    st_stream<NT>(atp(a.Fout + c * P, idu), fo);
    if constexpr (kOpCopy<T, S>) st_stream<NT>(atp(a.Fo_out + c * P, idu), (T)fo);
  } else {
    const float d = (float)df;
    if (!a.dead_rates) st_stream<NT>(atp(a.DF + c * P, idu), d);
    const float y = (float)a.Bdt * d - comp;
    const float t = fin + y;
    st_stream<NT>(atp(a.Fc + c * P, idu), (t - fin) - y);
    st_stream<NT>(atp(a.Fout + c * P, idu), t);
  }
}

}  // namespace wql
