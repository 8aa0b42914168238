// Two-block q-interface of the fused stage: hat variables, characteristic SAT
// forcing of both blocks, slip-rate update and stage-1 fault output, one
// thread per interface node.
//
// Follows, per node and in the same order of operations:
//   enforce_bound_iface_conditions (domain.f90:615-644): block L's evaluation
//     (its side 2) first, then block R's (its side 1);
//   Couple_Interface_x (CouplingForcing.f90:9-463): tractions n.sigma of both
//     faces (:107-140), rotation into (n, m, l) (Rotate_in_Local_Cordinates,
//     :466-522), hats, stage-1 outputs (:200-226), rotation back (:525-578),
//     SAT forcing u_bc / v_bc with S^T Z S and S^T Z^-1 S (:228-457);
//   Interface_RHS (Interface_Condition.f90:12-513): Welded_Interface (:515-590)
//     or Slip_Weakening_Friction (:745-916) with solve (:1095-1168), the
//     prestress and the linear slip-weakening strength of the problem taken
//     from per-node tables (built at wql_iface_create), or
//     rate_state_friction (:918-1067) with the regula falsi slip-rate solve
//     (:1120-1265) and the state rate (:420-444), or Linear_Friction
//     (:639-742); DS(1:3) += Q - P,
//     DS(4) += |slip rate|;
//   rupture_front (CouplingForcing.f90:580-600);
//   scale_rates_iface / update_fields_iface (iface.f90:58-85).
// The interface state is block L's copy (one slip-rate increment per stage,
// as each MPI rank holds its own interface).  All interface arithmetic is
// FP64, on FP64 copies of the coupled faces' frames and material.
#include <math.h>

#include "wql3d_internal.h"

namespace wql {
namespace {

// hats P (minus side) and Q (plus side) in (n, m, l) from the rotated grid
// values U (block L) and V (block R): velocities [0..2], tractions [3..5]
// T0: prestress (n, m, l); fs, fd, Dc: strength parameters of the node
__device__ __forceinline__ void hats(int coupling, const double U[6], const double V[6],
                                     double Z1s, double Z2s, double Z1p, double Z2p, double slip,
                                     const double T0[3], double fs, double fd, double Dc,
                                     double P[6], double Q[6]) {
  const double p_n = Z1p * U[0] - U[3], p_m = Z1s * U[1] - U[4], p_l = Z1s * U[2] - U[5];
  const double q_n = Z2p * V[0] + V[3], q_m = Z2s * V[1] + V[4], q_l = Z2s * V[2] + V[5];
  const double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
  const double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
  const double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
  const double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
  const double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
  if (coupling == WQL_IFACE_LOCKED) {
    P[0] = (q_n - phi_n) / Z2p; P[3] = phi_n;
    Q[0] = (p_n + phi_n) / Z1p; Q[3] = phi_n;
    P[2] = (q_l - phi_l) / Z2s; P[5] = phi_l;
    Q[2] = (p_l + phi_l) / Z1s; Q[5] = phi_l;
    P[1] = (q_m - phi_m) / Z2s; P[4] = phi_m;
    Q[1] = (p_m + phi_m) / Z1s; Q[4] = phi_m;
    return;
  }
  if (coupling == WQL_IFACE_LINEAR_FRICTION) {
    // Linear_Friction (Interface_Condition.f90:639-742), alpha in fd: normal
    // components as welded, shear tau = alpha/(alpha + eta_s) phi
    const double alpha = fd;
    P[0] = (q_n - phi_n) / Z2p; P[3] = phi_n;
    Q[0] = (p_n + phi_n) / Z1p; Q[3] = phi_n;
    const double Tm = alpha / (alpha + eta_s) * phi_m, Tl = alpha / (alpha + eta_s) * phi_l;
    P[4] = Tm; P[5] = Tl;
    Q[4] = Tm; Q[5] = Tl;
    Q[1] = (p_m + Tm) / Z1s + 1.0 / (alpha + eta_s) * phi_m;
    Q[2] = (p_l + Tl) / Z1s + 1.0 / (alpha + eta_s) * phi_l;
    P[1] = (q_m - Tm) / Z2s - 1.0 / (alpha + eta_s) * phi_m;
    P[2] = (q_l - Tl) / Z2s - 1.0 / (alpha + eta_s) * phi_l;
    return;
  }
  const double u_n_hat = (q_n - phi_n) / Z2p, Tu_n_hat = phi_n;
  const double v_n_hat = (p_n + phi_n) / Z1p, Tv_n_hat = phi_n;
  const double tau_lock =
      sqrt((T0[2] + phi_l) * (T0[2] + phi_l) + (T0[1] + phi_m) * (T0[1] + phi_m));
  const double sigma_n = fmax(0.0, -(T0[0] + phi_n));
  // linear slip weakening (Tau_strength)
  const double tau_str = sigma_n * (fs - (fs - fd) * fmin(slip, Dc) / Dc);
  double Tu_m, Tu_l, Tv_m, Tv_l, vv_m, vv_l;
  if (tau_lock > tau_str) {
    // sliding: |tau| = strength, slip rate V = (|phi| - strength) / eta_s
    const double phi_1 = T0[1] + phi_m, phi_2 = T0[2] + phi_l;
    const double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
    const double Vs = (Phi - tau_str) / eta_s;
    vv_m = phi_1 / (eta_s + tau_str / Vs);
    vv_l = phi_2 / (eta_s + tau_str / Vs);
    const double tau_m = phi_1 - eta_s * vv_m;
    const double tau_l = phi_2 - eta_s * vv_l;
    Tu_m = tau_m - T0[1];
    Tu_l = tau_l - T0[2];
    Tv_m = tau_m - T0[1];
    Tv_l = tau_l - T0[2];
  } else {
    Tu_m = phi_m; Tu_l = phi_l;
    Tv_m = phi_m; Tv_l = phi_l;
    vv_m = 0.0; vv_l = 0.0;
  }
  P[0] = u_n_hat;
  P[1] = (q_m - Tv_m) / Z2s - vv_m;
  P[2] = (q_l - Tv_l) / Z2s - vv_l;
  P[3] = Tu_n_hat; P[4] = Tu_m; P[5] = Tu_l;
  Q[0] = v_n_hat;
  Q[1] = (Tu_m + p_m) / Z1s + vv_m;
  Q[2] = (Tu_l + p_l) / Z1s + vv_l;
  Q[3] = Tv_n_hat; Q[4] = Tv_m; Q[5] = Tv_l;
}

// f of the rate-and-state slip-rate equation (Interface_Condition.f90:1241-1265)
__device__ __forceinline__ double rs_f(double V, double Phi, double eta, double sigma_n,
                                       double psi, double V0, double a) {
  return V + a * sigma_n / eta * asinh(0.5 * V / V0 * exp(psi / a)) - Phi / eta;
}

// Regula_Falsi (Interface_Condition.f90:1170-1239), step for step: bracket
// [0, Phi/eta], the reference's update rule and stopping test (|V - V_prev|
// < 1e-12 or 5000 iterations)
__device__ double rs_regula_falsi(double V, double Phi, double eta, double sigma_n, double psi,
                                  double V0, double a) {
  const double tol = 1.0e-12;
  double err = 1.0;
  double Vl = 0.0, Vr = Phi / eta;
  int k = 1;
  const int maxit = 5000;
  double fv = rs_f(V, Phi, eta, sigma_n, psi, V0, a);
  double fl = rs_f(Vl, Phi, eta, sigma_n, psi, V0, a);
  double fr = rs_f(Vr, Phi, eta, sigma_n, psi, V0, a);
  if (fabs(fv) <= tol) return V;
  if (fabs(fl) <= tol) return Vl;
  if (fabs(fr) <= tol) return Vr;
  while (k <= maxit && err >= tol) {
    const double V1 = V;
    if (fv * fl > tol) {
      Vl = V;
      V = Vr - (Vr - Vl) * fr / (fr - fl);
      fl = fv;
    } else if (fv * fr > tol) {
      Vr = V;
      V = Vr - (Vr - Vl) * fr / (fr - fl);
      fr = fv;
    }
    fv = rs_f(V, Phi, eta, sigma_n, psi, V0, a);
    err = fabs(V - V1);
    k = k + 1;
  }
  return V;
}

// rate_state_friction (Interface_Condition.f90:918-1067) with solve's
// rate-and-state branch (:1120-1146); returns the slip rate in *Vel
__device__ __forceinline__ void hats_rate_state(const double U[6], const double V[6], double Z1s,
                                                double Z2s, double Z1p, double Z2p,
                                                const double T0[3], double psi, double a,
                                                double V0, double P[6], double Q[6],
                                                double* Vel) {
  const double p_n = Z1p * U[0] - U[3], p_m = Z1s * U[1] - U[4], p_l = Z1s * U[2] - U[5];
  const double q_n = Z2p * V[0] + V[3], q_m = Z2s * V[1] + V[4], q_l = Z2s * V[2] + V[5];
  const double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
  const double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
  const double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
  const double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
  const double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
  const double sigma_n = fmax(0.0, -(T0[0] + phi_n));
  double vv_m = V[1] - U[1], vv_l = V[2] - U[2];
  const double phi_1 = T0[1] + phi_m, phi_2 = T0[2] + phi_l;
  const double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
  double Vs = sqrt(vv_m * vv_m + vv_l * vv_l);
  if (Vs > Phi) Vs = 0.5 * Phi / eta_s;
  Vs = rs_regula_falsi(Vs, Phi, eta_s, sigma_n, psi, V0, a);
  *Vel = Vs;
  const double fv = sigma_n * a * asinh(0.5 * Vs / V0 * exp(psi / a));
  vv_m = phi_1 / (eta_s + fv / Vs);
  vv_l = phi_2 / (eta_s + fv / Vs);
  const double tau_m = phi_1 - eta_s * vv_m, tau_l = phi_2 - eta_s * vv_l;
  const double Tm = tau_m - T0[1], Tl = tau_l - T0[2];
  P[0] = (q_n - phi_n) / Z2p;
  P[1] = (q_m - Tm) / Z2s - vv_m;
  P[2] = (q_l - Tl) / Z2s - vv_l;
  P[3] = phi_n; P[4] = Tm; P[5] = Tl;
  Q[0] = (p_n + phi_n) / Z1p;
  Q[1] = (Tm + p_m) / Z1s + vv_m;
  Q[2] = (Tl + p_l) / Z1s + vv_l;
  Q[3] = phi_n; Q[4] = Tm; Q[5] = Tl;
}

__device__ __forceinline__ void rot_in(const double n[3], const double m[3], const double l[3],
                                       double u[6]) {
  const double u1 = u[0], u2 = u[1], u3 = u[2], T1 = u[3], T2 = u[4], T3 = u[5];
  u[0] = n[0] * u1 + n[1] * u2 + n[2] * u3;
  u[1] = m[0] * u1 + m[1] * u2 + m[2] * u3;
  u[2] = l[0] * u1 + l[1] * u2 + l[2] * u3;
  u[3] = n[0] * T1 + n[1] * T2 + n[2] * T3;
  u[4] = m[0] * T1 + m[1] * T2 + m[2] * T3;
  u[5] = l[0] * T1 + l[1] * T2 + l[2] * T3;
}

__device__ __forceinline__ void rot_back(const double n[3], const double m[3], const double l[3],
                                         double u[6]) {
  const double u1 = u[0], u2 = u[1], u3 = u[2], T1 = u[3], T2 = u[4], T3 = u[5];
  u[0] = n[0] * u1 + m[0] * u2 + l[0] * u3;
  u[1] = n[1] * u1 + m[1] * u2 + l[1] * u3;
  u[2] = n[2] * u1 + m[2] * u2 + l[2] * u3;
  u[3] = n[0] * T1 + m[0] * T2 + l[0] * T3;
  u[4] = n[1] * T1 + m[1] * T2 + l[1] * T3;
  u[5] = n[2] * T1 + m[2] * T2 + l[2] * T3;
}

struct Side {
  double lam, mu, rho;
};

__device__ __forceinline__ bool is_rate_state(int coupling) {
  return coupling == WQL_IFACE_RATE_STATE_TPV101 || coupling == WQL_IFACE_RATE_STATE_TPV102 ||
         coupling == WQL_IFACE_RATE_STATE_RW;
}

// the state rate of problem 'rate-weakening' (Interface_Condition.f90:
// 438-444): -V/L0 (a asinh(V/(2 V0) e^(psi/a)) - Fss), Fss the steady state
// f0 - (b - a) ln(V/V0), flash-weakened towards 0.13 above V_w
__device__ __forceinline__ double rw_state_rate(double W, double Vel, double a, double Vw,
                                                double L0, double f0, double b, double V0) {
  double Fss = f0 - (b - a) * log(Vel / V0);
  const double r = Vel / Vw, r2 = r * r, r4 = r2 * r2;
  Fss = 0.13 + (Fss - 0.13) / pow(1.0 + r4 * r4, 0.125);
  return -(Vel / L0 * (a * asinh(Vel / (2.0 * V0) * exp(W / a)) - Fss));
}

// one evaluation (the calling block's frame and metric column): hats,
// slip-rate magnitude and the calling block's SAT forcing bc[9]
template <typename T, typename S>
__device__ __forceinline__ void evaluate(const IfaceArgs<T, S>& a, int ib, const double* fr,
                                         long long fi, long long NF, const double u[9],
                                         const double v[9], const Side& s1, const Side& s2,
                                         double slip, double psi, const double T0[3], double P[6],
                                         double Q[6], double* slip_rate, double* Vel,
                                         double bc[9]) {
  double l[3], m[3], n[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    l[k] = fr[k * NF + fi];
    m[k] = fr[(3 + k) * NF + fi];
    n[k] = fr[(6 + k) * NF + fi];
  }
  const double qx = fr[10 * NF + fi], qy = fr[11 * NF + fi], qz = fr[12 * NF + fi];
  double U[6], V[6];
  U[3] = n[0] * u[3] + n[1] * u[6] + n[2] * u[7];
  U[4] = n[0] * u[6] + n[1] * u[4] + n[2] * u[8];
  U[5] = n[0] * u[7] + n[1] * u[8] + n[2] * u[5];
  V[3] = n[0] * v[3] + n[1] * v[6] + n[2] * v[7];
  V[4] = n[0] * v[6] + n[1] * v[4] + n[2] * v[8];
  V[5] = n[0] * v[7] + n[1] * v[8] + n[2] * v[5];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    U[c] = u[c];
    V[c] = v[c];
  }
  rot_in(n, m, l, U);
  rot_in(n, m, l, V);
  const double c1p = sqrt((2.0 * s1.mu + s1.lam) / s1.rho), c1s = sqrt(s1.mu / s1.rho);
  const double c2p = sqrt((2.0 * s2.mu + s2.lam) / s2.rho), c2s = sqrt(s2.mu / s2.rho);
  const double Z1s = s1.rho * c1s, Z2s = s2.rho * c2s, Z1p = s1.rho * c1p, Z2p = s2.rho * c2p;
  *Vel = 0.0;
  if (is_rate_state(a.coupling)) {
    hats_rate_state(U, V, Z1s, Z2s, Z1p, Z2p, T0, psi, a.rs_a[fi], a.rs_V0, P, Q, Vel);
  } else {
    double fs = 0.0, fd = a.alpha, Dc = 1.0;
    if (a.coupling != WQL_IFACE_LOCKED && a.coupling != WQL_IFACE_LINEAR_FRICTION) {
      fs = a.fs[fi];
      fd = a.fd[fi];
      Dc = a.Dc[fi];
    }
    hats(a.coupling, U, V, Z1s, Z2s, Z1p, Z2p, slip, T0, fs, fd, Dc, P, Q);
  }
  *slip_rate = sqrt((Q[1] - P[1]) * (Q[1] - P[1]) + (Q[2] - P[2]) * (Q[2] - P[2]));
  double Pb[6], Qb[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    Pb[c] = P[c];
    Qb[c] = Q[c];
  }
  rot_back(n, m, l, U);
  rot_back(n, m, l, V);
  rot_back(n, m, l, Pb);
  rot_back(n, m, l, Qb);
  const double norm = sqrt(qx * qx + qy * qy + qz * qz);
  // rows n, m, l of the rotation
  const double* r[3] = {n, m, l};
  if (ib == 2) {
    // minus side: p = 1/2 (S^T Zp S v + T), forcing from (p - p_hat)
    double dp[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double mz[3];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        mz[j] = r[i][0] * r[j][0] * Z1p + r[i][1] * r[j][1] * Z1s + r[i][2] * r[j][2] * Z1s;
      const double ph = 0.5 * (mz[0] * Pb[0] + mz[1] * Pb[1] + mz[2] * Pb[2] + Pb[3 + i]);
      const double pp = 0.5 * (mz[0] * U[0] + mz[1] * U[1] + mz[2] * U[2] + U[3 + i]);
      dp[i] = pp - ph;
    }
    double A[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double mi[3];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        mi[j] = r[i][0] * r[j][0] * 1.0 / Z1p + r[i][1] * r[j][1] * 1.0 / Z1s +
                r[i][2] * r[j][2] * 1.0 / Z1s;
      A[i] = mi[0] * dp[0] + mi[1] * dp[1] + mi[2] * dp[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) bc[i] = norm / s1.rho * dp[i];
    const double l2m = 2.0 * s1.mu + s1.lam, lam = s1.lam, mu = s1.mu;
    bc[3] = l2m * qx * A[0] + lam * qy * A[1] + lam * qz * A[2];
    bc[4] = l2m * qy * A[1] + lam * qz * A[2] + lam * qx * A[0];
    bc[5] = l2m * qz * A[2] + lam * qy * A[1] + lam * qx * A[0];
    bc[6] = mu * (qy * A[0] + qx * A[1]);
    bc[7] = mu * (qz * A[0] + qx * A[2]);
    bc[8] = mu * (qz * A[1] + qy * A[2]);
  } else {
    // plus side: q = 1/2 (S^T Zq S v - T), forcing from (q - q_hat)
    double dq[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double nz[3];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        nz[j] = r[i][0] * r[j][0] * Z2p + r[i][1] * r[j][1] * Z2s + r[i][2] * r[j][2] * Z2s;
      const double qh = 0.5 * (nz[0] * Qb[0] + nz[1] * Qb[1] + nz[2] * Qb[2] - Qb[3 + i]);
      const double qm = 0.5 * (nz[0] * V[0] + nz[1] * V[1] + nz[2] * V[2] - V[3 + i]);
      dq[i] = qm - qh;
    }
    double B[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double qi[3];
#pragma unroll
      for (int j = 0; j < 3; ++j)
        qi[j] = r[i][0] * r[j][0] * 1.0 / Z2p + r[i][1] * r[j][1] * 1.0 / Z2s +
                r[i][2] * r[j][2] * 1.0 / Z2s;
      B[i] = qi[0] * dq[0] + qi[1] * dq[1] + qi[2] * dq[2];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) bc[i] = norm / s2.rho * dq[i];
    const double l2m = 2.0 * s2.mu + s2.lam, lam = s2.lam, mu = s2.mu;
    bc[3] = -l2m * qx * B[0] - lam * qy * B[1] - lam * qz * B[2];
    bc[4] = -l2m * qy * B[1] - lam * qz * B[2] - lam * qx * B[0];
    bc[5] = -l2m * qz * B[2] - lam * qy * B[1] - lam * qx * B[0];
    bc[6] = -mu * (qy * B[0] + qx * B[1]);
    bc[7] = -mu * (qz * B[0] + qx * B[2]);
    bc[8] = -mu * (qz * B[1] + qy * B[2]);
  }
}

// stage-1 outputs of one evaluation (CouplingForcing.f90:200-226,
// Interface_Condition.f90:494-503, rupture_front)
template <typename T, typename S>
__device__ __forceinline__ void stage1_output(const IfaceArgs<T, S>& a, const double P[6],
                                              const double Q[6], const double T0[3],
                                              double out[15], double* trup) {
  // out: Svel[3], Uhat[6], Vhat[6]
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    out[c] = Q[c] - P[c];
    out[3 + c] = Q[c] - P[c];
    out[6 + c] = P[3 + c] + T0[c];
    out[9 + c] = Q[c] - P[c];
    out[12 + c] = Q[3 + c] + T0[c];
  }
  if (sqrt(out[1] * out[1] + out[2] * out[2]) > 1e-3) *trup = fmin(*trup, a.t);
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_iface(const IfaceArgs<T, S> a) {
  const long long NF = (long long)a.ny * a.nz;
  const long long fi = (long long)blockIdx.x * 256 + threadIdx.x;
  if (fi >= NF) return;
  const int y = (int)(fi % a.ny), z = (int)(fi / a.ny);
  const long long iL = (long long)(a.nxL - 1) + (long long)a.nxL * (y + (long long)a.ny * z);
  const long long iR = (long long)a.nxR * (y + (long long)a.ny * z);  // block R's face is x = 0
  double u[9], v[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    u[c] = (double)a.FL[c * a.PL + iL];
    v[c] = (double)a.FR[c * a.PR + iR];
  }
  Side s1, s2;
  s1.lam = a.matL[fi]; s1.mu = a.matL[NF + fi]; s1.rho = a.matL[2 * NF + fi];
  s2.lam = a.matR[fi]; s2.mu = a.matR[NF + fi]; s2.rho = a.matR[2 * NF + fi];
  const double slip = a.S[3 * NF + fi];
  const bool rate_state = is_rate_state(a.coupling);
  const double psi = rate_state ? a.W[fi] : 0.0;
  // prestress of each side (zero for a welded interface, whose output
  // carries no prestress, Interface_Condition.f90:482-486, and for a linear
  // friction interface without tables)
  double T0L[3] = {0.0, 0.0, 0.0}, T0R[3] = {0.0, 0.0, 0.0};
  if (a.coupling != WQL_IFACE_LOCKED && a.T0L) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T0L[c] = a.T0L[c * NF + fi];
      T0R[c] = a.T0R[c * NF + fi];
    }
    if (a.coupling == WQL_IFACE_RATE_STATE_RW) {
      // Gaussian pulse 2 tau_pulse F(r) G(t) on T0_l, tau_pulse = -0.2429 T0_n,
      // G = 1 from t > 0 (Interface_Condition.f90:2408-2424)
      const double G = a.t > 0.0 ? 1.0 : 0.0;
      T0L[2] = T0L[2] + 2.0 * (-0.2429 * T0L[0]) * a.rs_nuc[fi] * G;
      T0R[2] = T0R[2] + 2.0 * (-0.2429 * T0R[0]) * a.rs_nuc[fi] * G;
    } else if (rate_state) {
      // nucleation 25 F(r) G(t) on T0_l (Interface_Condition.f90:2340-2347)
      const double t = a.t;
      double G = 0.0;
      if (t > 0.0 && t < 1.0) G = exp((t - 1.0) * (t - 1.0) / (t * (t - 2.0)));
      if (t >= 1.0) G = 1.0;
      T0L[2] = T0L[2] + a.rs_nuc[fi] * G;
      T0R[2] = T0R[2] + a.rs_nuc[fi] * G;
    }
  }
  double P[6], Q[6], srate, vel, bc[9];
  double trup = a.trup[fi];
  double out[15];
  // block L, side 2: SAT into L, the slip-rate increment
  evaluate(a, 2, a.frL, fi, NF, u, v, s1, s2, slip, psi, T0L, P, Q, &srate, &vel, bc);
  double dw = 0.0;
  if (a.coupling == WQL_IFACE_RATE_STATE_RW)
    dw = a.A * a.DW[fi] + rw_state_rate(psi, vel, a.rs_a[fi], a.rs_vw[fi], a.rs_L0, a.rs_f0,
                                        a.rs_b, a.rs_V0);
  else if (rate_state)   // the state rate of TPV101 / TPV102 (Interface_Condition.f90:431-436)
    dw = a.A * a.DW[fi] + (a.rs_b * a.rs_V0 / a.rs_L0 * exp(-(psi - a.rs_f0) / a.rs_b) -
                           vel * a.rs_b / a.rs_L0);
  double ds[4];
#pragma unroll
  for (int c = 0; c < 3; ++c) ds[c] = a.A * a.DS[c * NF + fi] + (Q[c] - P[c]);
  ds[3] = a.A * a.DS[3 * NF + fi] + srate;
  if (a.stage1) stage1_output(a, P, Q, T0L, out, &trup);
#pragma unroll
  for (int c = 0; c < 9; ++c) a.satL[c * NF + fi] = a.penL * bc[c];
  // block R, side 1: SAT into R (its hats are the same up to rounding)
  evaluate(a, 1, a.frR, fi, NF, u, v, s1, s2, slip, psi, T0R, P, Q, &srate, &vel, bc);
  if (a.stage1) stage1_output(a, P, Q, T0R, out, &trup);
#pragma unroll
  for (int c = 0; c < 9; ++c) a.satR[c * NF + fi] = a.penR * bc[c];
  if (a.stage1) {
    if (a.frame) {
      double* f = a.frame;
#pragma unroll
      for (int c = 0; c < 9; ++c) {
        f[c * NF + fi] = u[c];
        f[(9 + c) * NF + fi] = v[c];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) f[(18 + c) * NF + fi] = a.S[c * NF + fi];
#pragma unroll
      for (int c = 0; c < 15; ++c) f[(22 + c) * NF + fi] = out[c];
      f[37 * NF + fi] = trup;
      f[38 * NF + fi] = psi;  // the state at the start of the step, as I%W at stage 1
    }
    a.trup[fi] = trup;
  }
  // scale_rates_iface, then update_fields_iface after the blocks' stage
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    a.DS[c * NF + fi] = ds[c];
    a.S[c * NF + fi] = a.S[c * NF + fi] + a.Bdt * ds[c];
  }
  if (rate_state) {
    a.DW[fi] = dw;
    a.W[fi] = psi + a.Bdt * dw;
  }
}

}  // namespace

template <typename T, typename S>
hipError_t launch_iface(const IfaceArgs<T, S>& a, hipStream_t s) {
  const long long NF = (long long)a.ny * a.nz;
  hipLaunchKernelGGL((k_iface<T, S>), dim3((unsigned)((NF + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

template hipError_t launch_iface<double, double>(const IfaceArgs<double, double>&, hipStream_t);
template hipError_t launch_iface<float, double>(const IfaceArgs<float, double>&, hipStream_t);
template hipError_t launch_iface<float, float>(const IfaceArgs<float, float>&, hipStream_t);

}  // namespace wql
