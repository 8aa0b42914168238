/*
 * wql_oracle_iface.c -- CPU restatement (FP64) of the two-block interface of
 * WaveQLab3D-Q: the q-direction interface between the x-high face of block L
 * and the x-low face of block R: locked, TPV5 / TPV28 linear slip-weakening,
 * TPV101 / TPV102 or rate-weakening rate-and-state friction, or linear
 * friction.
 * TEST INFRASTRUCTURE ONLY (see wql_oracle.h).
 *
 * Follows, step by step and in the same order of operations:
 *   enforce_iface_conditions (domain.f90:615-680): one evaluation per block,
 *     block L first (its side 2), then block R (its side 1), both on the same
 *     interface object;
 *   Impose_Interface_Condition (RHS_Interior.f90:1771-1843): SAT forcing
 *     DF(face) -= tau0/hq * forcing of the calling block;
 *   Couple_Interface_x (CouplingForcing.f90:9-463): tractions, rotation to
 *     (n, m, l), hat variables, stage-1 outputs, rotation back, characteristic
 *     SAT forcing u_bc / v_bc;
 *   Interface_RHS (Interface_Condition.f90:12-513): hats by Welded_Interface
 *     (:515-590) or Slip_Weakening_Friction (:745-916) with solve (:1095-1168),
 *     Tau_strength TPV5 / TPV28 (:1283-1299, 1459-1475), prestress TPV5 /
 *     TPV28 (:1802-1830, 2049-2079), boxcar (:2438-2457); or
 *     rate_state_friction (:918-1067) with solve's rate-and-state branch
 *     (:1120-1146), Regula_Falsi and f (:1170-1265), prestress TPV101 /
 *     TPV102 (:2324-2378) and the state rate (:420-436); or the same
 *     solve for problem 'rate-weakening' (parameters :410-416, prestress
 *     :2387-2426 with initial_stress_tensor, initial_stress_condition.f90:
 *     184-191, state rate :438-444); or Linear_Friction (:639-742, alpha
 *     :360-366); DS(1:3) += Q - P,
 *     DS(4) += |slip rate|;
 *   rupture_front (CouplingForcing.f90:580-600);
 *   scale_rates_iface / update_fields_iface (iface.f90:58-85).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "wql_oracle.h"

static double boxcar(double x, double w) {
  const double tol = 100.0 * 2.220446049250313e-16;
  if (-w + tol < x && x < w - tol) return 1.0;
  if (fabs(-w - x) <= tol || fabs(x - w) <= tol) return 0.5;
  return 0.0;
}

/* prestress, problem TPV5 (Interface_Condition.f90:1821-1830) */
static void prestress_tpv5(double y, double z, double *T0n, double *T0m, double *T0l) {
  *T0n = -120.0;
  *T0m = 0.0;
  double l = 70.0;
  l = l + 8.0 * boxcar(y - 7.5, 1.5) * boxcar(z + 7.5, 1.5);
  l = l + 11.6 * boxcar(y - 7.5, 1.5) * boxcar(z, 1.5);
  l = l - 8.0 * boxcar(y - 7.5, 1.5) * boxcar(z - 7.5, 1.5);
  *T0l = l;
}

/* prestress, problem TPV28 (Interface_Condition.f90:2049-2079): the traction
 * of the uniform initial stress tensor s0 = (xx, yy, zz, xy, xz, yz)
 * (initial_stress_condition.f90:78-84) on the face, in (n, m, l), plus the
 * nucleation stress on T0_l */
static void prestress_tpv28(const double *s0, const double *n, const double *m, const double *l,
                            double y, double z, double *T0n, double *T0m, double *T0l) {
  const double pi = 3.141592653589793;
  const double sxx = s0[0], syy = s0[1], szz = s0[2], sxy = s0[3], sxz = s0[4], syz = s0[5];
  const double Tx = n[0] * sxx + n[1] * sxy + n[2] * sxz;
  const double Ty = n[0] * sxy + n[1] * syy + n[2] * syz;
  const double Tz = n[0] * sxz + n[1] * syz + n[2] * szz;
  *T0n = n[0] * Tx + n[1] * Ty + n[2] * Tz;
  *T0m = m[0] * Tx + m[1] * Ty + m[2] * Tz;
  *T0l = l[0] * Tx + l[1] * Ty + l[2] * Tz;
  const double r = sqrt((y - 7.5) * (y - 7.5) + z * z);
  double tau_nuke;
  if (r <= 1.4)
    tau_nuke = 11.60;
  else if (r >= 1.4 && r <= 2.0)
    tau_nuke = 5.8 * (1.0 + cos(pi * (r - 1.4) / 0.6));
  else
    tau_nuke = 0.0;
  *T0l = *T0l + tau_nuke;
}

/* prestress, problems TPV101 / TPV102 (Interface_Condition.f90:2324-2378):
 * the traction of the initial stress tensor (-120, 0, -120, 0, 75, 0)
 * (initial_stress_condition.f90:114-120) on the face, in (n, m, l), plus the
 * nucleation stress 25 F(r) G(t) on T0_l, F(r) = exp(r^2 / (r^2 - 9)) for
 * r < 3 km around (y, z) = (7.5, 0), G ramping from 0 to 1 over 1 s */
static void prestress_tpv101(const double *n, const double *m, const double *l, double y,
                             double z, double t, double *T0n, double *T0m, double *T0l) {
  const double sxx = -120.0, syy = 0.0, szz = -120.0, sxy = 0.0, sxz = 75.0, syz = 0.0;
  const double Tx = n[0] * sxx + n[1] * sxy + n[2] * sxz;
  const double Ty = n[0] * sxy + n[1] * syy + n[2] * syz;
  const double Tz = n[0] * sxz + n[1] * syz + n[2] * szz;
  *T0n = n[0] * Tx + n[1] * Ty + n[2] * Tz;
  *T0m = m[0] * Tx + m[1] * Ty + m[2] * Tz;
  *T0l = l[0] * Tx + l[1] * Ty + l[2] * Tz;
  const double r = sqrt((y - 7.5) * (y - 7.5) + z * z);
  double F = 0.0;
  if (r < 3.0) F = exp(r * r / (r * r - 9.0));
  double G = 0.0;
  if (t > 0.0 && t < 1.0) G = exp((t - 1.0) * (t - 1.0) / (t * (t - 2.0)));
  if (t >= 1.0) G = 1.0;
  *T0l = *T0l + 25.0 * F * G;
}

/* prestress, problem 'rate-weakening' (Interface_Condition.f90:2387-2426):
 * the traction of initial_stress_tensor('rate-weakening')
 * (initial_stress_condition.f90:184-191: psi = 50 degrees with that file's
 * pi = 3.14159265359; sxx = -126, sxz = 42, szz from psi, syy their mean)
 * on the face, in (n, m, l), plus the Gaussian pulse 2 tau_pulse F G on T0_l,
 * tau_pulse = -0.2429 T0_n, F = exp(-(r/r0)^2/2), r0 = 3*1.5*0.3 around
 * (y, z) = (8, 0), G = 1 for t > 0 */
static void prestress_rw(const double *n, const double *m, const double *l, double y, double z,
                         double t, double *T0n, double *T0m, double *T0l) {
  const double pi = 3.14159265359;
  const double psi = 50.0 / 180.0 * pi;
  double s0[6];
  s0[0] = -126.0;
  s0[4] = 42.0;
  s0[2] = (1.0 - 2.0 * s0[4] / (s0[0] * tan(2.0 * psi))) * s0[0];
  s0[1] = 0.5 * (s0[0] + s0[2]);
  s0[3] = 0.0;
  s0[5] = 0.0;
  const double Tx = n[0] * s0[0] + n[1] * s0[3] + n[2] * s0[4];
  const double Ty = n[0] * s0[3] + n[1] * s0[1] + n[2] * s0[5];
  const double Tz = n[0] * s0[4] + n[1] * s0[5] + n[2] * s0[2];
  *T0n = n[0] * Tx + n[1] * Ty + n[2] * Tz;
  *T0m = m[0] * Tx + m[1] * Ty + m[2] * Tz;
  *T0l = l[0] * Tx + l[1] * Ty + l[2] * Tz;
  const double tau_pulse = -0.2429 * *T0n;
  const double r0 = 3.0 * 1.5 * 0.3;
  const double y0 = 8.0, z0 = 0.0;
  const double r = sqrt((y - y0) * (y - y0) + (z + z0) * (z + z0));
  const double F = exp(-((r / r0) * (r / r0)) / 2.0);
  double G = 0.0;
  if (t > 0.0) G = 1.0;
  *T0l = *T0l + 2.0 * tau_pulse * F * G;
}

void orc_rw_prestress(const double *n, const double *m, const double *l, double y, double z,
                      double t, double T0[3]) {
  prestress_rw(n, m, l, y, z, t, &T0[0], &T0[1], &T0[2]);
}

static int is_rate_state(int coupling) {
  return coupling == ORC_IFACE_RATE_STATE_TPV101 || coupling == ORC_IFACE_RATE_STATE_TPV102 ||
         coupling == ORC_IFACE_RATE_STATE_RW;
}

/* prestress of the friction problems in (n, m, l) at time t */
static void prestress(const orc_iface *I, const double *n, const double *m, const double *l,
                      double y, double z, double t, double T0[3]) {
  if (I->coupling == ORC_IFACE_SLIP_WEAKENING_TPV28)
    prestress_tpv28(I->s0, n, m, l, y, z, &T0[0], &T0[1], &T0[2]);
  else if (I->coupling == ORC_IFACE_RATE_STATE_RW)
    prestress_rw(n, m, l, y, z, t, &T0[0], &T0[1], &T0[2]);
  else if (I->coupling == ORC_IFACE_LINEAR_FRICTION)
    T0[0] = T0[1] = T0[2] = 0.0;
  else if (is_rate_state(I->coupling))
    prestress_tpv101(n, m, l, y, z, t, &T0[0], &T0[1], &T0[2]);
  else
    prestress_tpv5(y, z, &T0[0], &T0[1], &T0[2]);
}

/* smooth_boxcar (Interface_Condition.f90:2460-2476) */
static double smooth_boxcar(double x, double w, double wr) {
  if (-w <= x && x <= w) return 1.0;
  if (fabs(x) > w && fabs(x) < w + wr)
    return 0.5 * (1.0 + tanh(wr / (fabs(x) - w - wr) + wr / (fabs(x) - w)));
  return 0.0;
}

double orc_rs_a(double y, double z) {
  return 0.008 + 0.008 * (1.0 - smooth_boxcar(y, 15.0, 3.0) * smooth_boxcar(z, 15.0, 3.0));
}

double orc_rw_a(double y, double z) {
  return 0.016 + 0.016 * (1.0 - smooth_boxcar(y, 20.0, 2.5) * smooth_boxcar(z, 20.0, 2.5));
}

double orc_rw_vw(double y, double z) {
  return 0.17 + 5.0 * (1.0 - smooth_boxcar(y, 20.0, 2.5) * smooth_boxcar(z, 20.0, 2.5));
}

/* the state rate of problem 'rate-weakening' (Interface_Condition.f90:438-444),
 * L0 = 2*0.2572, f0 = 0.7, b = 0.02, V0 = 1e-6 (:410-416); (V/V_w)^8 as
 * three squarings */
double orc_rw_state_rate(double W, double Vel, double a, double Vw) {
  const double L0 = 2.0 * 0.2572, f0 = 0.7, b = 0.02, V0 = 1.0e-6;
  double Fss = f0 - (b - a) * log(Vel / V0);
  const double r = Vel / Vw, r2 = r * r, r4 = r2 * r2;
  Fss = 0.13 + (Fss - 0.13) / pow(1.0 + r4 * r4, 0.125);
  return -(Vel / L0 * (a * asinh(Vel / (2.0 * V0) * exp(W / a)) - Fss));
}

void orc_welded_hats(const double U[6], const double V[6], double Z1s, double Z2s, double Z1p,
                     double Z2p, double P[6], double Q[6]) {
  const double p_n = Z1p * U[0] - U[3], p_m = Z1s * U[1] - U[4], p_l = Z1s * U[2] - U[5];
  const double q_n = Z2p * V[0] + V[3], q_m = Z2s * V[1] + V[4], q_l = Z2s * V[2] + V[5];
  const double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
  const double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
  const double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
  const double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
  const double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
  P[0] = (q_n - phi_n) / Z2p; P[3] = phi_n;
  Q[0] = (p_n + phi_n) / Z1p; Q[3] = phi_n;
  P[2] = (q_l - phi_l) / Z2s; P[5] = phi_l;
  Q[2] = (p_l + phi_l) / Z1s; Q[5] = phi_l;
  P[1] = (q_m - phi_m) / Z2s; P[4] = phi_m;
  Q[1] = (p_m + phi_m) / Z1s; Q[4] = phi_m;
}

/* Linear_Friction (Interface_Condition.f90:639-742): the normal components
 * as the welded interface, the shear components from tau = alpha v and the
 * characteristics; the reference does not store u_out(2) (the minus side's m
 * velocity), which is u_m_hat of the same derivation here */
void orc_linear_friction_hats(const double U[6], const double V[6], double Z1s, double Z2s,
                              double Z1p, double Z2p, double alpha, double P[6], double Q[6]) {
  const double p_n = Z1p * U[0] - U[3], p_m = Z1s * U[1] - U[4], p_l = Z1s * U[2] - U[5];
  const double q_n = Z2p * V[0] + V[3], q_m = Z2s * V[1] + V[4], q_l = Z2s * V[2] + V[5];
  const double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
  const double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
  const double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
  const double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
  const double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
  const double u_n_hat = (q_n - phi_n) / Z2p, Tu_n_hat = phi_n;
  const double v_n_hat = (p_n + phi_n) / Z1p, Tv_n_hat = phi_n;
  const double Tu_m_hat = alpha / (alpha + eta_s) * phi_m;
  const double Tu_l_hat = alpha / (alpha + eta_s) * phi_l;
  const double Tv_m_hat = alpha / (alpha + eta_s) * phi_m;
  const double Tv_l_hat = alpha / (alpha + eta_s) * phi_l;
  const double v_m_hat = (p_m + Tu_m_hat) / (Z1s) + 1.0 / (alpha + eta_s) * phi_m;
  const double v_l_hat = (p_l + Tu_l_hat) / (Z1s) + 1.0 / (alpha + eta_s) * phi_l;
  const double u_m_hat = (q_m - Tv_m_hat) / (Z2s) - 1.0 / (alpha + eta_s) * phi_m;
  const double u_l_hat = (q_l - Tv_l_hat) / (Z2s) - 1.0 / (alpha + eta_s) * phi_l;
  P[0] = u_n_hat; P[1] = u_m_hat; P[2] = u_l_hat;
  P[3] = Tu_n_hat; P[4] = Tu_m_hat; P[5] = Tu_l_hat;
  Q[0] = v_n_hat; Q[1] = v_m_hat; Q[2] = v_l_hat;
  Q[3] = Tv_n_hat; Q[4] = Tv_m_hat; Q[5] = Tv_l_hat;
}

/* f (Interface_Condition.f90:1241-1265) */
static double rs_f(double V, double Phi, double eta, double sigma_n, double psi, double V0,
                   double a) {
  return V + a * sigma_n / eta * asinh(0.5 * V / V0 * exp(psi / a)) - Phi / eta;
}

double orc_rs_regula_falsi(double V, double Phi, double eta, double sigma_n, double psi,
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

/* the state rate of TPV101 / TPV102 (Interface_Condition.f90:431-436) */
double orc_rs_state_rate(double W, double Vel, double L0, double f0, double b, double V0) {
  return b * V0 / L0 * exp(-(W - f0) / b) - Vel * b / L0;
}

/* Tau_strength, problems TPV5 (Interface_Condition.f90:1283-1299) and TPV28
 * (:1459-1475): linear slip weakening with a high-strength barrier outside
 * the rupture area */
static double tau_strength(int coupling, double sigma_n, double D, double y, double z) {
  double fs = 0.677, fd, Dc = 0.4;
  if (coupling == ORC_IFACE_SLIP_WEAKENING_TPV28) {
    fd = 0.373;
    fs = fs + 1.0e10 * (1.0 - boxcar(y, 15.0) * boxcar(z, 18.0));
  } else {
    fd = 0.525;
    fs = fs + 1.0e10 * (1.0 - boxcar(y, 15.0) * boxcar(z, 15.0));
  }
  const double f = fs - (fs - fd) * fmin(D, Dc) / Dc;
  return sigma_n * f;
}

/* hats P (left) and Q (right) in local (n, m, l) coordinates from the rotated
 * grid values U, V; returns the slip-rate magnitude of the friction solve */
static void hats(const orc_iface *I, const double *n, const double *m, const double *l,
                 const double U[6], const double V[6], double Z1s, double Z2s, double Z1p,
                 double Z2p, double slip, double psi, double y, double z, double t, double P[6],
                 double Q[6], double *Vel) {
  const int coupling = I->coupling;
  *Vel = 0.0;
  const double p_n = Z1p * U[0] - U[3], p_m = Z1s * U[1] - U[4], p_l = Z1s * U[2] - U[5];
  const double q_n = Z2p * V[0] + V[3], q_m = Z2s * V[1] + V[4], q_l = Z2s * V[2] + V[5];
  if (coupling == ORC_IFACE_LINEAR_FRICTION) {
    orc_linear_friction_hats(U, V, Z1s, Z2s, Z1p, Z2p, I->alpha, P, Q);
    return;
  }
  if (coupling == ORC_IFACE_LOCKED) {
    /* Welded_Interface (Interface_Condition.f90:515-590) */
    const double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
    const double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
    const double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
    const double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
    const double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
    P[0] = (q_n - phi_n) / Z2p; P[3] = phi_n;
    Q[0] = (p_n + phi_n) / Z1p; Q[3] = phi_n;
    P[2] = (q_l - phi_l) / Z2s; P[5] = phi_l;
    Q[2] = (p_l + phi_l) / Z1s; Q[5] = phi_l;
    P[1] = (q_m - phi_m) / Z2s; P[4] = phi_m;
    Q[1] = (p_m + phi_m) / Z1s; Q[4] = phi_m;
    return;
  }
  if (is_rate_state(coupling)) {
    /* rate_state_friction (Interface_Condition.f90:918-1067) with solve's
     * rate-and-state branch (:1120-1146); parameters of TPV101 / TPV102
     * (:386-400) or rate-weakening (:410-416) */
    const double V0 = 1.0e-6;
    const double a = coupling == ORC_IFACE_RATE_STATE_RW ? orc_rw_a(y, z) : orc_rs_a(y, z);
    const double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
    const double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
    double T0[3];
    prestress(I, n, m, l, y, z, t, T0);
    const double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
    const double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
    const double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
    const double u_n_hat = (q_n - phi_n) / Z2p, Tu_n_hat = phi_n;
    const double v_n_hat = (p_n + phi_n) / Z1p, Tv_n_hat = phi_n;
    const double sigma_n = fmax(0.0, -(T0[0] + phi_n));
    double vv_m = V[1] - U[1], vv_l = V[2] - U[2];
    const double phi_1 = T0[1] + phi_m, phi_2 = T0[2] + phi_l;
    const double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
    double Vs = sqrt(vv_m * vv_m + vv_l * vv_l);
    if (Vs > Phi) Vs = 0.5 * Phi / eta_s;
    Vs = orc_rs_regula_falsi(Vs, Phi, eta_s, sigma_n, psi, V0, a);
    *Vel = Vs;
    const double fv = sigma_n * a * asinh(0.5 * Vs / V0 * exp(psi / a));
    vv_m = phi_1 / (eta_s + fv / Vs);
    vv_l = phi_2 / (eta_s + fv / Vs);
    const double tau_m = phi_1 - eta_s * vv_m, tau_l = phi_2 - eta_s * vv_l;
    const double Tu_m_hat = tau_m - T0[1], Tu_l_hat = tau_l - T0[2];
    const double Tv_m_hat = tau_m - T0[1], Tv_l_hat = tau_l - T0[2];
    P[0] = u_n_hat;
    P[1] = (q_m - Tv_m_hat) / Z2s - vv_m;
    P[2] = (q_l - Tv_l_hat) / Z2s - vv_l;
    P[3] = Tu_n_hat; P[4] = Tu_m_hat; P[5] = Tu_l_hat;
    Q[0] = v_n_hat;
    Q[1] = (Tu_m_hat + p_m) / Z1s + vv_m;
    Q[2] = (Tu_l_hat + p_l) / Z1s + vv_l;
    Q[3] = Tv_n_hat; Q[4] = Tv_m_hat; Q[5] = Tv_l_hat;
    return;
  }
  /* Slip_Weakening_Friction (Interface_Condition.f90:745-916) */
  const double eta_s = 1.0 / (1.0 / Z1s + 1.0 / Z2s);
  const double eta_p = 1.0 / (1.0 / Z1p + 1.0 / Z2p);
  double T0[3];
  prestress(I, n, m, l, y, z, t, T0);
  const double T0n = T0[0], T0m = T0[1], T0l = T0[2];
  const double phi_n = eta_p * (q_n / Z2p - p_n / Z1p);
  const double phi_m = eta_s * (q_m / Z2s - p_m / Z1s);
  const double phi_l = eta_s * (q_l / Z2s - p_l / Z1s);
  const double u_n_hat = (q_n - phi_n) / Z2p, Tu_n_hat = phi_n;
  const double v_n_hat = (p_n + phi_n) / Z1p, Tv_n_hat = phi_n;
  const double tau_lock = sqrt((T0l + phi_l) * (T0l + phi_l) + (T0m + phi_m) * (T0m + phi_m));
  const double sigma_n = fmax(0.0, -(T0n + phi_n));
  const double tau_str = tau_strength(coupling, sigma_n, slip, y, z);
  double Tu_m_hat, Tu_l_hat, Tv_m_hat, Tv_l_hat, vv_m, vv_l;
  if (tau_lock > tau_str) {
    /* solve, slip-weakening branch (Interface_Condition.f90:1143-1160) */
    const double phi_1 = T0m + phi_m, phi_2 = T0l + phi_l;
    const double Phi = sqrt(phi_1 * phi_1 + phi_2 * phi_2);
    const double V = (Phi - tau_str) / eta_s;
    vv_m = phi_1 / (eta_s + tau_str / V);
    vv_l = phi_2 / (eta_s + tau_str / V);
    const double tau_m = phi_1 - eta_s * vv_m;
    const double tau_l = phi_2 - eta_s * vv_l;
    Tu_m_hat = tau_m - T0m;
    Tu_l_hat = tau_l - T0l;
    Tv_m_hat = tau_m - T0m;
    Tv_l_hat = tau_l - T0l;
  } else {
    Tu_m_hat = phi_m;
    Tu_l_hat = phi_l;
    Tv_m_hat = phi_m;
    Tv_l_hat = phi_l;
    vv_m = 0.0;
    vv_l = 0.0;
  }
  const double v_m_hat = (Tu_m_hat + p_m) / Z1s + vv_m;
  const double v_l_hat = (Tu_l_hat + p_l) / Z1s + vv_l;
  const double u_m_hat = (q_m - Tv_m_hat) / Z2s - vv_m;
  const double u_l_hat = (q_l - Tv_l_hat) / Z2s - vv_l;
  P[0] = u_n_hat; P[1] = u_m_hat; P[2] = u_l_hat;
  P[3] = Tu_n_hat; P[4] = Tu_m_hat; P[5] = Tu_l_hat;
  Q[0] = v_n_hat; Q[1] = v_m_hat; Q[2] = v_l_hat;
  Q[3] = Tv_n_hat; Q[4] = Tv_m_hat; Q[5] = Tv_l_hat;
}

/* to local (n, m, l): Rotate_in_Local_Cordinates (CouplingForcing.f90:466-522) */
static void rot_in(const double *n, const double *m, const double *l, double *u) {
  const double u1 = u[0], u2 = u[1], u3 = u[2], T1 = u[3], T2 = u[4], T3 = u[5];
  u[0] = n[0] * u1 + n[1] * u2 + n[2] * u3;
  u[1] = m[0] * u1 + m[1] * u2 + m[2] * u3;
  u[2] = l[0] * u1 + l[1] * u2 + l[2] * u3;
  u[3] = n[0] * T1 + n[1] * T2 + n[2] * T3;
  u[4] = m[0] * T1 + m[1] * T2 + m[2] * T3;
  u[5] = l[0] * T1 + l[1] * T2 + l[2] * T3;
}

/* back to x, y, z: Rotate_Back_Physical_Cordinates (CouplingForcing.f90:525-578) */
static void rot_back(const double *n, const double *m, const double *l, double *u) {
  const double u1 = u[0], u2 = u[1], u3 = u[2], T1 = u[3], T2 = u[4], T3 = u[5];
  u[0] = n[0] * u1 + m[0] * u2 + l[0] * u3;
  u[1] = n[1] * u1 + m[1] * u2 + l[1] * u3;
  u[2] = n[2] * u1 + m[2] * u2 + l[2] * u3;
  u[3] = n[0] * T1 + m[0] * T2 + l[0] * T3;
  u[4] = n[1] * T1 + m[1] * T2 + l[1] * T3;
  u[5] = n[2] * T1 + m[2] * T2 + l[2] * T3;
}

/* S^T Z S and S^T Z^-1 S as written in CouplingForcing.f90:316-341 */
static void char_mats(const double *n, const double *m, const double *l, double Zp, double Zs,
                      double Mz[3][3], double Mi[3][3]) {
  const double *r[3] = {n, m, l};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      /* entry (a,b): r_a(1) r_b(1) Zp + r_a(2) r_b(2) Zs + r_a(3) r_b(3) Zs */
      Mz[a][b] = r[a][0] * r[b][0] * Zp + r[a][1] * r[b][1] * Zs + r[a][2] * r[b][2] * Zs;
      Mi[a][b] = r[a][0] * r[b][0] * 1.0 / Zp + r[a][1] * r[b][1] * 1.0 / Zs +
                 r[a][2] * r[b][2] * 1.0 / Zs;
    }
}

static void eval_side(orc_iface *I, orc_block *L, orc_block *R, int ib, double t, int stage) {
  orc_block *B = ib == 2 ? L : R;
  const int ny = I->ny, nz = I->nz, NF = ny * nz;
  const int fL = 1, fR = 0;            /* Rx face of L, Lx face of R */
  const int fB = ib == 2 ? fL : fR;
  const long long NL = (long long)L->nx * L->ny * L->nz, NR = (long long)R->nx * R->ny * R->nz;
  const long long NB = ib == 2 ? NL : NR;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) {
      const int fi = y + ny * z;
      const long long iL = (L->nx - 1) + (long long)L->nx * (y + (long long)L->ny * z);
      const long long iR = 0 + (long long)R->nx * (y + (long long)R->ny * z);
      const long long iB = ib == 2 ? iL : iR;
      double n[3], m[3], l[3];
      for (int k = 0; k < 3; ++k) {
        n[k] = B->nn[fB][k * NF + fi];
        m[k] = B->nm[fB][k * NF + fi];
        l[k] = B->nl[fB][k * NF + fi];
      }
      const double *u = L->F, *v = R->F;
      /* tractions and velocities in x, y, z (CouplingForcing.f90:107-140) */
      double U[6], V[6];
      U[3] = n[0] * u[3 * NL + iL] + n[1] * u[6 * NL + iL] + n[2] * u[7 * NL + iL];
      U[4] = n[0] * u[6 * NL + iL] + n[1] * u[4 * NL + iL] + n[2] * u[8 * NL + iL];
      U[5] = n[0] * u[7 * NL + iL] + n[1] * u[8 * NL + iL] + n[2] * u[5 * NL + iL];
      V[3] = n[0] * v[3 * NR + iR] + n[1] * v[6 * NR + iR] + n[2] * v[7 * NR + iR];
      V[4] = n[0] * v[6 * NR + iR] + n[1] * v[4 * NR + iR] + n[2] * v[8 * NR + iR];
      V[5] = n[0] * v[7 * NR + iR] + n[1] * v[8 * NR + iR] + n[2] * v[5 * NR + iR];
      for (int c = 0; c < 3; ++c) {
        U[c] = u[c * NL + iL];
        V[c] = v[c * NR + iR];
      }
      rot_in(n, m, l, U);
      rot_in(n, m, l, V);
      /* Interface_RHS: materials of both sides at the node */
      const double lam1 = L->M[iL], mu1 = L->M[NL + iL], rho1 = L->M[2 * NL + iL];
      const double lam2 = R->M[iR], mu2 = R->M[NR + iR], rho2 = R->M[2 * NR + iR];
      const double c1p = sqrt((2.0 * mu1 + lam1) / rho1), c1s = sqrt(mu1 / rho1);
      const double c2p = sqrt((2.0 * mu2 + lam2) / rho2), c2s = sqrt(mu2 / rho2);
      const double Z1s = rho1 * c1s, Z2s = rho2 * c2s, Z1p = rho1 * c1p, Z2p = rho2 * c2p;
      double P[6], Q[6], Vel;
      hats(I, n, m, l, U, V, Z1s, Z2s, Z1p, Z2p, I->S[3 * NF + fi], I->W[fi], I->yf[fi],
           I->zf[fi], t, P, Q, &Vel);
      const double slip_rate = sqrt((Q[1] - P[1]) * (Q[1] - P[1]) + (Q[2] - P[2]) * (Q[2] - P[2]));
      /* the interface state belongs to block L's side: one DS increment per
       * stage, as when each block keeps its own interface copy */
      if (ib == 2) {
        for (int c = 0; c < 3; ++c) I->DS[c * NF + fi] = I->DS[c * NF + fi] + Q[c] - P[c];
        I->DS[3 * NF + fi] = I->DS[3 * NF + fi] + slip_rate;
        if (I->coupling == ORC_IFACE_RATE_STATE_RW)
          I->DW[fi] = I->DW[fi] + orc_rw_state_rate(I->W[fi], Vel, orc_rw_a(I->yf[fi], I->zf[fi]),
                                                    orc_rw_vw(I->yf[fi], I->zf[fi]));
        else if (is_rate_state(I->coupling))
          I->DW[fi] = I->DW[fi] + orc_rs_state_rate(I->W[fi], Vel, 0.02, 0.6, 0.012, 1.0e-6);
      }
      if (stage == 1) {
        double T0[3] = {0.0, 0.0, 0.0};
        if (I->coupling != ORC_IFACE_LOCKED) prestress(I, n, m, l, I->yf[fi], I->zf[fi], t, T0);
        for (int c = 0; c < 3; ++c) {
          I->Uhat[(3 + c) * NF + fi] = P[3 + c] + T0[c];
          I->Vhat[(3 + c) * NF + fi] = Q[3 + c] + T0[c];
          /* CouplingForcing.f90:210-219: slip velocity in local coordinates */
          I->Svel[c * NF + fi] = Q[c] - P[c];
          I->Uhat[c * NF + fi] = Q[c] - P[c];
          I->Vhat[c * NF + fi] = Q[c] - P[c];
        }
        if (sqrt(I->Svel[1 * NF + fi] * I->Svel[1 * NF + fi] +
                 I->Svel[2 * NF + fi] * I->Svel[2 * NF + fi]) > 1e-3)
          I->trup[fi] = fmin(I->trup[fi], t);
      }
      rot_back(n, m, l, U);
      rot_back(n, m, l, V);
      rot_back(n, m, l, P);
      rot_back(n, m, l, Q);
      /* characteristic SAT forcing (CouplingForcing.f90:228-457); both
       * blocks use the metric column of the calling block */
      const double qx = B->metx[iB], qy = B->mety[iB], qz = B->metz[iB];
      const double norm = sqrt(qx * qx + qy * qy + qz * qz);
      (void)NB;
      double Mz[3][3], Pi[3][3], Nz[3][3], Qi[3][3];
      char_mats(n, m, l, Z1p, Z1s, Mz, Pi);
      char_mats(n, m, l, Z2p, Z2s, Nz, Qi);
      double dp[3], dq[3];
      for (int a = 0; a < 3; ++a) {
        const double ph = 0.5 * (Mz[a][0] * P[0] + Mz[a][1] * P[1] + Mz[a][2] * P[2] + P[3 + a]);
        const double pp = 0.5 * (Mz[a][0] * U[0] + Mz[a][1] * U[1] + Mz[a][2] * U[2] + U[3 + a]);
        const double qh = 0.5 * (Nz[a][0] * Q[0] + Nz[a][1] * Q[1] + Nz[a][2] * Q[2] - Q[3 + a]);
        const double qm = 0.5 * (Nz[a][0] * V[0] + Nz[a][1] * V[1] + Nz[a][2] * V[2] - V[3 + a]);
        dp[a] = pp - ph;
        dq[a] = qm - qh;
      }
      double bc[9];
      if (ib == 2) {
        double A[3];
        for (int a = 0; a < 3; ++a) A[a] = Pi[a][0] * dp[0] + Pi[a][1] * dp[1] + Pi[a][2] * dp[2];
        for (int a = 0; a < 3; ++a) bc[a] = norm / rho1 * dp[a];
        bc[3] = (2.0 * mu1 + lam1) * qx * A[0] + lam1 * qy * A[1] + lam1 * qz * A[2];
        bc[4] = (2.0 * mu1 + lam1) * qy * A[1] + lam1 * qz * A[2] + lam1 * qx * A[0];
        bc[5] = (2.0 * mu1 + lam1) * qz * A[2] + lam1 * qy * A[1] + lam1 * qx * A[0];
        bc[6] = mu1 * (qy * A[0] + qx * A[1]);
        bc[7] = mu1 * (qz * A[0] + qx * A[2]);
        bc[8] = mu1 * (qz * A[1] + qy * A[2]);
      } else {
        double Bq[3];
        for (int a = 0; a < 3; ++a) Bq[a] = Qi[a][0] * dq[0] + Qi[a][1] * dq[1] + Qi[a][2] * dq[2];
        for (int a = 0; a < 3; ++a) bc[a] = norm / rho2 * dq[a];
        bc[3] = -(2.0 * mu2 + lam2) * qx * Bq[0] - lam2 * qy * Bq[1] - lam2 * qz * Bq[2];
        bc[4] = -(2.0 * mu2 + lam2) * qy * Bq[1] - lam2 * qz * Bq[2] - lam2 * qx * Bq[0];
        bc[5] = -(2.0 * mu2 + lam2) * qz * Bq[2] - lam2 * qy * Bq[1] - lam2 * qx * Bq[0];
        bc[6] = -mu2 * (qy * Bq[0] + qx * Bq[1]);
        bc[7] = -mu2 * (qz * Bq[0] + qx * Bq[2]);
        bc[8] = -mu2 * (qz * Bq[1] + qy * Bq[2]);
      }
      /* Impose_Interface_Condition: DF(face) -= tau0/hq * forcing */
      const double pen = B->tau0 / B->hq;
      for (int c = 0; c < 9; ++c) B->DF[c * NB + iB] = B->DF[c * NB + iB] - pen * bc[c];
    }
}

void orc_iface_enforce(orc_iface *I, orc_block *L, orc_block *R, double t, int stage, int side) {
  eval_side(I, L, R, side == 2 ? 2 : 1, t, stage);
}

void orc_iface_scale(orc_iface *I, double A) {
  const int NF = I->ny * I->nz;
  for (int k = 0; k < 4 * NF; ++k) I->DS[k] = A * I->DS[k];
  for (int k = 0; k < NF; ++k) I->DW[k] = A * I->DW[k];
}

void orc_iface_update(orc_iface *I, double inc) {
  const int NF = I->ny * I->nz;
  for (int k = 0; k < 4 * NF; ++k) I->S[k] = I->S[k] + inc * I->DS[k];
  for (int k = 0; k < NF; ++k) I->W[k] = I->W[k] + inc * I->DW[k];
}
