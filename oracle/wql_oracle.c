/*
 * wql_oracle.c -- CPU restatement (FP64, scalar C) of the WaveQLab3D-Q
 * per-block RK time step, used ONLY as the correctness checker for the HIP
 * product path (tests/, __graft_entry__.smoke(), bench.py cpu_baseline leg).
 *
 * Each routine follows the reference Fortran routine it names, including the
 * order of floating-point operations where that is practical, so that the
 * restatement reproduces the reference's printed known-answer values.
 * Citations are to files of aaimran/waveqlab3d_Q_GPU (src/...).
 *
 * Built with -O2 -ffp-contract=off (no FMA contraction), like the reference's
 * cpu-release x86-64 build, which has no FMA in its baseline ISA.
 *
 * Parity status of each part (how it is pinned) is documented in the test
 * files tests/test_oracle_*.py:
 *   orc_rk_coefficients       pinned: rationals + order-4 convergence
 *   orc_sbp_closure / JJU_x6  pinned: SBP identity H D + (H D)^T = B, H11,
 *                             polynomial exactness, mirror antisymmetry
 *   orc_metrics               pinned: exact for affine maps, J = det
 *   orc_set_rates (+SAT, src) pinned: reference KAT max|field| 41^3 (16 digits),
 *                             constant-state zero rates, energy decay
 *   Q4/Q8                     pinned: closed-form relaxation, modulus formula
 *   PML                       pinned: d=0 reduces to elastic; slab/profile
 *                             quirks and the feedback terms (F = 0, curved
 *                             mesh) checked against closed forms; auxiliary
 *                             source terms exact for cubic fields on a
 *                             non-diagonal affine curvilinear map
 *   orc_source_time           pinned: unit integrals (Gaussian, cosine bell);
 *                             Ricker = -g''/(2 pi^2 f0^2), zero mean, spectral
 *                             peak at f0
 */
#include "wql_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Node loops whose iterations write only their own node (the RHS, the RK
 * vector updates) run on OpenMP threads; each node's arithmetic is the same
 * on any thread, so the results do not depend on the thread count. */
int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

#define IDX(b, x, y, z) ((long)((x) - 1) + (long)(b)->nx * ((long)((y) - 1) + (long)(b)->ny * (long)((z) - 1)))

/* ---------------------------------------------------------------------------
 * SBP-6 left closure rows, terms in the reference's order
 * (JU_xJU_yJU_z6.f90:74-182; metrics.f90:30-97).  col is 1-based.
 * The kind-less literal 0.013435342414630 (row 4, column 7) is taken as
 * double, as under the cpu-release -fdefault-real-8 build.
 * ------------------------------------------------------------------------- */
typedef struct { int n; double c[8]; int col[8]; } sbp_row_t;
static const sbp_row_t LROW[6] = {
  {6, {0.036577936277544, 0.104488069284042, -0.450398306578272, -0.141512858744873,
       2.033378678700676, -1.582533518939116}, {6, 5, 4, 3, 2, 1}},
  {5, {-0.014903449191300, -0.069112065532624, 0.258816087376832, 0.287258622978251,
       -0.462059195631158}, {6, 5, 4, 3, 1}},
  {5, {-0.018129342917256, -0.022902190275815, 0.606235523609147, -0.636451095137907,
       0.071247104721830}, {6, 5, 4, 2, 1}},
  {6, {0.013435342414630, -0.051642265516119, 0.520262285050482, -0.306681191361148,
       -0.290087484386815, 0.114713313798970}, {7, 6, 5, 3, 2, 1}},
  {7, {0.018281071473911, -0.164529643265203, 0.769199413962647, -0.707905442575989,
       0.015764336127392, 0.105400944933782, -0.036210680656541}, {8, 7, 6, 4, 3, 2, 1}},
  {8, {0.016437980868017, -0.147941827812150, 0.739709139060752, -0.691649024426814,
       0.063183694641876, 0.011220896474665, 0.020437334208704, -0.011398193015050},
   {9, 8, 7, 5, 4, 3, 2, 1}},
};
/* central row (JU_xJU_yJU_z6.f90:53-59): offsets -3..+3 without 0 */
static const double CC[6] = {-1.0 / 60.0, 0.15, -0.75, 0.75, -0.15, 1.0 / 60.0};
static const int CO[6] = {-3, -2, -1, 1, 2, 3};

void orc_sbp_closure(double *t) {
  memset(t, 0, sizeof(double) * 54);
  for (int i = 0; i < 6; ++i)
    for (int k = 0; k < LROW[i].n; ++k) t[i * 9 + LROW[i].col[k] - 1] = LROW[i].c[k];
}

/* Build the list of (coef, col) terms of D1 row i of an n-point operator,
 * in the reference's term order. Returns count. */
static int d1_terms(int i, int n, double *c, int *col) {
  if (i >= 7 && i <= n - 6) {
    for (int k = 0; k < 6; ++k) { c[k] = CC[k]; col[k] = i + CO[k]; }
    return 6;
  }
  if (i <= 6) {
    const sbp_row_t *r = &LROW[i - 1];
    for (int k = 0; k < r->n; ++k) { c[k] = r->c[k]; col[k] = r->col[k]; }
    return r->n;
  }
  /* right closure: row n+1-i' mirrored, negated coefficients
   * (JU_xJU_yJU_z6.f90:185-299) */
  {
    const sbp_row_t *r = &LROW[n - i];
    for (int k = 0; k < r->n; ++k) { c[k] = -r->c[k]; col[k] = n + 1 - r->col[k]; }
    return r->n;
  }
}

/* D1 (row terms of JJU_x6 / D6x) applied to a 1-D array u[n]: du = (1/h) D u */
void orc_apply_d1(int n, double h, const double *u, double *du) {
  double c[8]; int col[8];
  for (int i = 1; i <= n; ++i) {
    int nt = d1_terms(i, n, c, col);
    double s = 0.0;
    for (int k = 0; k < nt; ++k) s = (k == 0) ? c[k] * u[col[k] - 1] : s + c[k] * u[col[k] - 1];
    du[i - 1] = (1.0 / h) * s;
  }
}

/* ---------------------------------------------------------------------------
 * RK coefficients, Kennedy-Carpenter (5,4) 2N-storage (time_step.f90:53-62)
 * ------------------------------------------------------------------------- */
void orc_rk_coefficients(double *A, double *B, double *C) {
  A[0] = 0.0;
  A[1] = -567301805773.0 / 1357537059087.0;
  A[2] = -2404267990393.0 / 2016746695238.0;
  A[3] = -3550918686646.0 / 2091501179385.0;
  A[4] = -1275806237668.0 / 842570457699.0;
  B[0] = 1432997174477.0 / 9575080441755.0;
  B[1] = 5161836677717.0 / 13612068292357.0;
  B[2] = 1720146321549.0 / 2090206949498.0;
  B[3] = 3134564353537.0 / 4481467310338.0;
  B[4] = 2277821191437.0 / 14882151754819.0;
  C[0] = 0.0;
  C[1] = 1432997174477.0 / 9575080441755.0;
  C[2] = 2526269341429.0 / 6820363962896.0;
  C[3] = 2006345519317.0 / 3224310063776.0;
  C[4] = 2802321613138.0 / 2924317926251.0;
  C[5] = 1.0;
}

/* ---------------------------------------------------------------------------
 * derivatives() (grid.f90:398-471) with D6x/D6y/D6z (metrics.f90:16-488)
 * ------------------------------------------------------------------------- */
static void d6(const double *X, int nx, int ny, int nz, int axis, double h,
               int x, int y, int z, double out[3]) {
  int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  int i = axis == 0 ? x : (axis == 1 ? y : z);
  double c[8]; int col[8];
  int nt = d1_terms(i, n, c, col);
  long N = (long)nx * ny * nz;
  for (int m = 0; m < 3; ++m) {
    double s = 0.0;
    for (int k = 0; k < nt; ++k) {
      int xx = x, yy = y, zz = z;
      if (axis == 0) xx = col[k]; else if (axis == 1) yy = col[k]; else zz = col[k];
      double v = X[m * N + (long)(xx - 1) + (long)nx * ((yy - 1) + (long)ny * (zz - 1))];
      s = (k == 0) ? c[k] * v : s + c[k] * v;
    }
    out[m] = (1.0 / h) * s;
  }
}

void orc_metrics(int nx, int ny, int nz, const double *X,
                 double *metx, double *mety, double *metz, double *J) {
  long N = (long)nx * ny * nz;
  double hq = 1.0 / (double)(nx - 1), hr = 1.0 / (double)(ny - 1), hs = 1.0 / (double)(nz - 1);
  for (int s = 1; s <= nz; ++s)
    for (int r = 1; r <= ny; ++r)
      for (int q = 1; q <= nx; ++q) {
        double Xq[3], Xr[3], Xs[3];
        d6(X, nx, ny, nz, 0, hq, q, r, s, Xq);
        d6(X, nx, ny, nz, 1, hr, q, r, s, Xr);
        d6(X, nx, ny, nz, 2, hs, q, r, s, Xs);
        long id = (long)(q - 1) + (long)nx * ((r - 1) + (long)ny * (s - 1));
        double j = Xq[0] * (Xr[1] * Xs[2] - Xr[2] * Xs[1])
                 - Xq[1] * (Xr[0] * Xs[2] - Xr[2] * Xs[0])
                 + Xq[2] * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
        J[id] = j;
        double ij = 1.0 / j;
        metx[0 * N + id] = ij * (Xr[1] * Xs[2] - Xr[2] * Xs[1]);
        metx[1 * N + id] = ij * (Xq[2] * Xs[1] - Xq[1] * Xs[2]);
        metx[2 * N + id] = ij * (Xq[1] * Xr[2] - Xq[2] * Xr[1]);
        mety[0 * N + id] = ij * (Xr[2] * Xs[0] - Xr[0] * Xs[2]);
        mety[1 * N + id] = ij * (Xq[0] * Xs[2] - Xq[2] * Xs[0]);
        mety[2 * N + id] = ij * (Xq[2] * Xr[0] - Xq[0] * Xr[2]);
        metz[0 * N + id] = ij * (Xr[0] * Xs[1] - Xr[1] * Xs[0]);
        metz[1 * N + id] = ij * (Xq[1] * Xs[0] - Xq[0] * Xs[1]);
        metz[2 * N + id] = ij * (Xq[0] * Xr[1] - Xq[1] * Xr[0]);
      }
}

/* ---------------------------------------------------------------------------
 * gen_orth_vectors / Gram_Schmidt (unit_normals.f90:232-284) and
 * local_orth_vectors_{q,r,s} (unit_normals.f90:8-168)
 * ------------------------------------------------------------------------- */
static void gram_schmidt(const double *y, double *z) {
  double a = y[0] * z[0] + y[1] * z[1] + y[2] * z[2];
  for (int k = 0; k < 3; ++k) z[k] = z[k] - a * y[k];
  double s = 1.0 / sqrt(pow(z[0], 2.0) + pow(z[1], 2.0) + pow(z[2], 2.0));
  for (int k = 0; k < 3; ++k) z[k] = s * z[k];
}

static void gen_orth(double *l, double *m, const double *n) {
  const double tol = 1.0e-12;
  m[0] = 0.0; m[1] = 1.0; m[2] = 0.0;
  double d1 = sqrt(pow(n[0] - m[0], 2.0) + pow(n[1] - m[1], 2.0) + pow(n[2] - m[2], 2.0));
  double d2 = sqrt(pow(n[0] + m[0], 2.0) + pow(n[1] + m[1], 2.0) + pow(n[2] + m[2], 2.0));
  if (!(d1 >= tol && d2 >= tol)) { m[0] = 0.0; m[1] = 0.0; m[2] = 1.0; }
  gram_schmidt(n, m);
  l[0] = n[1] * m[2] - n[2] * m[1];
  l[1] = n[2] * m[0] - n[0] * m[2];
  l[2] = n[0] * m[1] - n[1] * m[0];
}

void orc_normals(orc_block *b) {
  long N = (long)b->nx * b->ny * b->nz;
  for (int f = 0; f < 6; ++f) {
    if (!b->nn[f]) continue;
    int axis = f / 2;
    int fixed = (f % 2 == 0) ? 1 : (axis == 0 ? b->nx : (axis == 1 ? b->ny : b->nz));
    int n1 = axis == 0 ? b->ny : b->nx;
    int n2 = axis == 2 ? b->ny : b->nz;
    long NF = (long)n1 * n2;
    for (int j = 1; j <= n2; ++j)
      for (int i = 1; i <= n1; ++i) {
        int x, y, z;
        if (axis == 0) { x = fixed; y = i; z = j; }
        else if (axis == 1) { x = i; y = fixed; z = j; }
        else { x = i; y = j; z = fixed; }
        long id = IDX(b, x, y, z);
        double qx = b->metx[axis * N + id], qy = b->mety[axis * N + id], qz = b->metz[axis * N + id];
        double nrm = sqrt(pow(qx, 2.0) + pow(qy, 2.0) + pow(qz, 2.0));
        double n[3] = {qx / nrm, qy / nrm, qz / nrm}, l[3], m[3];
        gen_orth(l, m, n);
        long fid = (long)(i - 1) + (long)n1 * (j - 1);
        for (int k = 0; k < 3; ++k) {
          b->nl[f][k * NF + fid] = l[k];
          b->nm[f][k * NF + fid] = m[k];
          b->nn[f][k * NF + fid] = n[k];
        }
      }
  }
}

/* ---------------------------------------------------------------------------
 * Q4 / Q8 fixed coefficients (anelastic_q4_model.f90:97-111,
 * anelastic_q8_model.f90:100-116) and modulus correction
 * (material.f90:258-280 Q4, 425-453 Q8)
 * ------------------------------------------------------------------------- */
void orc_q_coefficients(int nmech, double fmin, double fmax, double *tau, double *weight) {
  const double pi = 3.141592653589793;
  double taumin = 1.0 / (2.0 * pi * fmax);
  double taumax = 1.0 / (2.0 * pi * fmin);
  for (int k = 1; k <= nmech; ++k)
    tau[k - 1] = exp(log(taumin) + (2.0 * k - 1.0) / (2.0 * nmech) * log(taumax / taumin));
  if (nmech == 8) {
    const double w[8] = {1.50589707, 0.0, 0.52793567, 0.53065494,
                         0.32862132, 0.64375916, 0.0, 1.32751442};
    memcpy(weight, w, sizeof w);
  } else {
    const double w[4] = {1.549360, 0.804277, 0.887718, 1.464160};
    memcpy(weight, w, sizeof w);
  }
}

int orc_q_modulus_correction(orc_block *b, double fref) {
  const double pi = 3.141592653589793;
  long N = (long)b->nx * b->ny * b->nz;
  double wref = 2.0 * pi * fref;
  for (long id = 0; id < N; ++id) {
    double vS = 0.0, vP = 0.0;
    for (int l = 0; l < b->nmech; ++l) {
      double dS = (wref * wref * (b->tau[l] * b->tau[l]) + 1.0) * (1.0 / b->Qs_inv[id]);
      double dP = (wref * wref * (b->tau[l] * b->tau[l]) + 1.0) * (1.0 / b->Qp_inv[id]);
      vS = vS + b->weight[l] / dS;
      vP = vP + b->weight[l] / dP;
    }
    if (vS >= 1.0 || vP >= 1.0) return -1;
    double lam = b->M[id], mu = b->M[N + id], rho = b->M[2 * N + id];
    double vs = sqrt(mu / rho);
    double vp = sqrt((lam + 2.0 * mu) / rho);
    double muS = rho * (vs * vs) / (1.0 - vS);
    double muP = rho * (vp * vp) / (1.0 - vP);
    vs = sqrt(muS / rho);
    vp = sqrt(muP / rho);
    b->M[N + id] = (vs * vs) * rho;
    b->M[id] = (vp * vp) * rho - 2.0 * b->M[N + id];
  }
  return 0;
}

/* ---------------------------------------------------------------------------
 * scale_rates_interior / update_fields_interior (fields.f90:51-263)
 * ------------------------------------------------------------------------- */
static long pml_slab_contains(const orc_block *b, int f, int x, int y, int z) {
  int n = b->npml[f];
  switch (f) {
    case 0: return x <= n;
    case 1: return x >= b->nx - n + 1;
    case 2: return y <= n;
    case 3: return y >= b->ny - n + 1;
    case 4: return z <= n;
    default: return z >= b->nz - n + 1;
  }
}

void orc_scale_rates(orc_block *b, double A) {
  long N = (long)b->nx * b->ny * b->nz;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 9 * N; ++i) b->DF[i] = A * b->DF[i];
  if (b->nmech) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 6L * b->nmech * N; ++i) b->Deta[i] = A * b->Deta[i];
  }
  for (int f = 0; f < 6; ++f) {
    if (!b->pml[f]) continue;
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= b->nz; ++z)
      for (int y = 1; y <= b->ny; ++y)
        for (int x = 1; x <= b->nx; ++x)
          if (pml_slab_contains(b, f, x, y, z))
            for (int c = 0; c < 9; ++c) {
              long i = c * N + IDX(b, x, y, z);
              b->pDQ[f][i] = A * b->pDQ[f][i];
            }
  }
}

void orc_update_fields(orc_block *b, double inc) {
  long N = (long)b->nx * b->ny * b->nz;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 9 * N; ++i) b->F[i] = b->F[i] + inc * b->DF[i];
  if (b->nmech) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 6L * b->nmech * N; ++i) b->eta[i] = b->eta[i] + inc * b->Deta[i];
  }
  for (int f = 0; f < 6; ++f) {
    if (!b->pml[f]) continue;
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= b->nz; ++z)
      for (int y = 1; y <= b->ny; ++y)
        for (int x = 1; x <= b->nx; ++x)
          if (pml_slab_contains(b, f, x, y, z))
            for (int c = 0; c < 9; ++c) {
              long i = c * N + IDX(b, x, y, z);
              b->pQ[f][i] = b->pQ[f][i] + inc * b->pDQ[f][i];
            }
  }
}

/* ---------------------------------------------------------------------------
 * JJU_x6 (JU_xJU_yJU_z6.f90:12-809): SBP derivative at one point for one
 * metric column array X (3 components: q, r, s directions)
 * ------------------------------------------------------------------------- */
static void jju_x6(const orc_block *b, int x, int y, int z, const double *X, double Ju[9]) {
  long N = (long)b->nx * b->ny * b->nz;
  const double *F = b->F, *J = b->J;
  long id0 = IDX(b, x, y, z);
  double part[3][9];
  int pos[3] = {x, y, z};
  int nn[3] = {b->nx, b->ny, b->nz};
  for (int a = 0; a < 3; ++a) {
    double c[8]; int col[8];
    int nt = d1_terms(pos[a], nn[a], c, col);
    double v[3] = {0, 0, 0}, s[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < nt; ++k) {
      int p[3] = {x, y, z};
      p[a] = col[k];
      long id = IDX(b, p[0], p[1], p[2]);
      for (int cc = 0; cc < 3; ++cc) {
        double t = c[k] * F[cc * N + id];
        v[cc] = (k == 0) ? t : v[cc] + t;
      }
      double jx = J[id], xx = X[a * N + id];
      for (int cc = 0; cc < 6; ++cc) {
        double t = ((c[k] * F[(3 + cc) * N + id]) * jx) * xx;
        s[cc] = (k == 0) ? t : s[cc] + t;
      }
    }
    double xpt = X[a * N + id0];
    for (int cc = 0; cc < 3; ++cc) part[a][cc] = xpt * v[cc];
    for (int cc = 0; cc < 6; ++cc) part[a][3 + cc] = s[cc];
  }
  double ih[3] = {1.0 / b->hq, 1.0 / b->hr, 1.0 / b->hs};
  for (int cc = 0; cc < 9; ++cc)
    Ju[cc] = ih[0] * part[0][cc] + ih[1] * part[1][cc] + ih[2] * part[2][cc];
}

/* interior extents: ix/fx etc. (JU_xJU_yJU_z6.f90:16158-16173) */
static void interior_extents(const orc_block *b, int *ix, int *fx, int *iy, int *fy, int *iz, int *fz) {
  *ix = *iy = *iz = *fx = *fy = *fz = 6;
  if (b->pml[4]) *iz = b->npml[4] > 6 ? b->npml[4] : 6;
  if (b->pml[5]) *fz = b->npml[5] > 6 ? b->npml[5] : 6;
  if (b->pml[2]) *iy = b->npml[2] > 6 ? b->npml[2] : 6;
  if (b->pml[3]) *fy = b->npml[3] > 6 ? b->npml[3] : 6;
  if (b->pml[0]) *ix = b->npml[0] > 6 ? b->npml[0] : 6;
  if (b->pml[1]) *fx = b->npml[1] > 6 ? b->npml[1] : 6;
}

/* memory-variable sums and rates shared by the interior and point routines
 * (JU_xJU_yJU_z6.f90:16335-16376; RHS_Interior.f90:2748-2793) */
static double eta_sum(const orc_block *b, int c6, long id) {
  long N = (long)b->nx * b->ny * b->nz;
  double s = b->eta[((long)c6 * b->nmech + 0) * N + id];
  for (int i = 1; i < b->nmech; ++i) s = s + b->eta[((long)c6 * b->nmech + i) * N + id];
  return s;
}

static void q_rates(orc_block *b, long id, double exx, double eyy, double ezz, double tr,
                    double gxy, double gxz, double gyz) {
  long N = (long)b->nx * b->ny * b->nz;
  double lam = b->M[id], mu = b->M[N + id];
  double Qs = b->Qs_inv[id], Qp = b->Qp_inv[id];
  double e[3] = {exx, eyy, ezz}, g[3] = {gxy, gxz, gyz};
  for (int i = 0; i < b->nmech; ++i) {
    double w = b->weight[i], ta = b->tau[i];
    for (int c = 0; c < 3; ++c) {
      long k = ((long)c * b->nmech + i) * N + id;
#if ORC_PLANT == 1
      /* planted mistake for the tests that must catch it: lambda in place of
       * lambda + 2 mu in the Qp term */
      b->Deta[k] = b->Deta[k] + ((((w * 2.0 * mu * Qs) * e[c]
                   + (w * (lam * Qp - 2.0 * mu * Qs)) * tr)) - b->eta[k]) / ta;
#else
      b->Deta[k] = b->Deta[k] + ((((w * 2.0 * mu * Qs) * e[c]
                   + (w * ((lam + 2.0 * mu) * Qp - 2.0 * mu * Qs)) * tr)) - b->eta[k]) / ta;
#endif
    }
    for (int c = 0; c < 3; ++c) {
      long k = ((long)(3 + c) * b->nmech + i) * N + id;
#if ORC_PLANT == 4
      /* planted mistake: 2 mu in place of mu in the shear source */
      b->Deta[k] = b->Deta[k] + ((w * 2.0 * mu * Qs * g[c] - b->eta[k]) / ta);
#else
      b->Deta[k] = b->Deta[k] + ((w * mu * Qs * g[c] - b->eta[k]) / ta);
#endif
    }
  }
}

/* RHS_Center -> JJU_x6_interior (JU_xJU_yJU_z6.f90:16088-16640) */
static void rhs_center(orc_block *b) {
  long N = (long)b->nx * b->ny * b->nz;
  int ix, fx, iy, fy, iz, fz;
  interior_extents(b, &ix, &fx, &iy, &fy, &iz, &fz);
  const double c3[3] = {0.75, -0.15, 1.0 / 60.0};
  double cx[3], cy[3], cz[3];
  for (int s = 0; s < 3; ++s) { cx[s] = c3[s] / b->hq; cy[s] = c3[s] / b->hr; cz[s] = c3[s] / b->hs; }
  const double *F = b->F, *J = b->J;
  const double *met[3] = {b->metx, b->mety, b->metz};
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = iz + 1; z <= b->nz - fz; ++z)
    for (int y = iy + 1; y <= b->ny - fy; ++y)
      for (int x = ix + 1; x <= b->nx - fx; ++x) {
        long id = IDX(b, x, y, z);
        double D[3][9];
        memset(D, 0, sizeof D);
        for (int s = 1; s <= 3; ++s) {
          long xp = IDX(b, x + s, y, z), xm = IDX(b, x - s, y, z);
          long yp = IDX(b, x, y + s, z), ym = IDX(b, x, y - s, z);
          long zp = IDX(b, x, y, z + s), zm = IDX(b, x, y, z - s);
          if (b->curvilinear) {
            for (int m = 0; m < 3; ++m) {
              const double *X = met[m];
              for (int c = 0; c < 3; ++c)
                D[m][c] = D[m][c]
                  + X[0 * N + id] * cx[s - 1] * (F[c * N + xp] - F[c * N + xm])
                  + X[1 * N + id] * cy[s - 1] * (F[c * N + yp] - F[c * N + ym])
                  + X[2 * N + id] * cz[s - 1] * (F[c * N + zp] - F[c * N + zm]);
              for (int c = 3; c < 9; ++c)
                D[m][c] = D[m][c]
                  + cx[s - 1] * (F[c * N + xp] * (J[xp] * X[0 * N + xp]) - F[c * N + xm] * (J[xm] * X[0 * N + xm]))
                  + cy[s - 1] * (F[c * N + yp] * (J[yp] * X[1 * N + yp]) - F[c * N + ym] * (J[ym] * X[1 * N + ym]))
                  + cz[s - 1] * (F[c * N + zp] * (J[zp] * X[2 * N + zp]) - F[c * N + zm] * (J[zm] * X[2 * N + zm]));
            }
          } else {
            for (int c = 0; c < 9; ++c) {
              D[0][c] = D[0][c] + b->metx[0 * N + id] * cx[s - 1] * (F[c * N + xp] - F[c * N + xm]);
              D[1][c] = D[1][c] + b->mety[1 * N + id] * cy[s - 1] * (F[c * N + yp] - F[c * N + ym]);
              D[2][c] = D[2][c] + b->metz[2 * N + id] * cz[s - 1] * (F[c * N + zp] - F[c * N + zm]);
            }
          }
        }
        double lam = b->M[id], mu = b->M[N + id], rho = b->M[2 * N + id];
        double a1 = b->curvilinear ? 1.0 / (rho * J[id]) : 1.0 / rho;
        double a2 = 2.0 * mu + lam, a3 = lam, a4 = mu;
        double *DF = b->DF;
        DF[0 * N + id] = DF[0 * N + id] + a1 * (D[0][3] + D[1][6] + D[2][7]);
        DF[1 * N + id] = DF[1 * N + id] + a1 * (D[0][6] + D[1][4] + D[2][8]);
        DF[2 * N + id] = DF[2 * N + id] + a1 * (D[0][7] + D[1][8] + D[2][5]);
        DF[3 * N + id] = DF[3 * N + id] + a2 * D[0][0] + a3 * (D[1][1] + D[2][2]);
        DF[4 * N + id] = DF[4 * N + id] + a2 * D[1][1] + a3 * (D[0][0] + D[2][2]);
        DF[5 * N + id] = DF[5 * N + id] + a2 * D[2][2] + a3 * (D[0][0] + D[1][1]);
        DF[6 * N + id] = DF[6 * N + id] + a4 * (D[1][0] + D[0][1]);
        DF[7 * N + id] = DF[7 * N + id] + a4 * (D[2][0] + D[0][2]);
        DF[8 * N + id] = DF[8 * N + id] + a4 * (D[2][1] + D[1][2]);
        if (b->nmech) {
          for (int c6 = 0; c6 < 6; ++c6) DF[(3 + c6) * N + id] = DF[(3 + c6) * N + id] - eta_sum(b, c6, id);
          double tr = D[0][0] + D[1][1] + D[2][2];
          q_rates(b, id, D[0][0], D[1][1], D[2][2], tr,
                  D[1][0] + D[0][1], D[2][0] + D[0][2], D[2][1] + D[1][2]);
        }
      }
}

/* PML damping at a node for face f (RHS_Interior.f90:1913,1977,2037,2096,2156,2215) */
static double pml_d(const orc_block *b, int f, int x, int y, int z) {
  const double Rf = 1.0e-3;
  const double d0 = 6.0 / (2.0 * 1.0) * log(1.0 / Rf);
  int np = b->npml[f];
  switch (f) {
    case 0: return d0 * fabs((double)(x - (1 + np)) / (double)np);
    case 1: return d0 * fabs((double)(x - (b->nx - np)) / (double)np);
    case 2: return d0 * fabs((double)(y - (1 + np)) / (double)np);
    case 3: return d0 * fabs((double)(y - (b->ny - np)) / (double)np);
    case 4: return d0 * fabs((double)(z - np) / (double)np);  /* z-low quirk */
    default: return d0 * fabs((double)(z - (b->nz - np)) / (double)np);
  }
}

/* PML (RHS_Interior.f90:1846-2265): face order 5,6,3,4,1,2 */
static void pml_point(orc_block *b, double DU[9], const double Ux[9], const double Uy[9],
                      const double Uz[9], int x, int y, int z, double rhoJ_inv, double lambda2mu) {
  long N = (long)b->nx * b->ny * b->nz;
  long id = IDX(b, x, y, z);
  double lam = b->M[id], mu = b->M[N + id];
  const double cfs = 1.0;
  static const int order[6] = {4, 5, 2, 3, 0, 1};
  for (int o = 0; o < 6; ++o) {
    int f = order[o];
    if (!b->pml[f] || !pml_slab_contains(b, f, x, y, z)) continue;
    double d = pml_d(b, f, x, y, z);
    double *Q = b->pQ[f], *DQ = b->pDQ[f];
    double q[9];
    for (int c = 0; c < 9; ++c) q[c] = Q[c * N + id];
    double Jp = b->J[id];
#if ORC_PLANT == 2
    /* planted mistake: the PML feedback with the opposite sign */
    d = -d;
#endif
    DU[0] = DU[0] - d * Jp * q[0] * rhoJ_inv;
    DU[1] = DU[1] - d * Jp * q[1] * rhoJ_inv;
    DU[2] = DU[2] - d * Jp * q[2] * rhoJ_inv;
    DU[3] = DU[3] - d * (lambda2mu * q[3] + lam * q[4] + lam * q[5]);
    DU[4] = DU[4] - d * (lam * q[3] + lambda2mu * q[4] + lam * q[5]);
    DU[5] = DU[5] - d * (lam * q[3] + lam * q[4] + lambda2mu * q[5]);
    DU[6] = DU[6] - mu * d * q[6];
    DU[7] = DU[7] - mu * d * q[7];
    DU[8] = DU[8] - mu * d * q[8];
    double src[9];
    int axis = f / 2;
    if (axis == 0) {
      double s[9] = {Ux[3] / Jp, Ux[6] / Jp, Ux[7] / Jp, Ux[0], 0.0, 0.0, Ux[1], Ux[2], 0.0};
      memcpy(src, s, sizeof s);
    } else if (axis == 1) {
      double s[9] = {Uy[6] / Jp, Uy[4] / Jp, Uy[8] / Jp, 0.0, Uy[1], 0.0, Uy[0], 0.0, Uy[2]};
      memcpy(src, s, sizeof s);
    } else {
      double s[9] = {Uz[7] / Jp, Uz[8] / Jp, Uz[5] / Jp, 0.0, 0.0, Uz[2], 0.0, Uz[0], Uz[1]};
      memcpy(src, s, sizeof s);
    }
#if ORC_PLANT == 2
    d = -d;
#endif
#if ORC_PLANT == 3
    /* planted mistake: the auxiliary source with the opposite sign */
    for (int c = 0; c < 9; ++c) src[c] = -src[c];
#endif
    for (int c = 0; c < 9; ++c) DQ[c * N + id] = DQ[c * N + id] + src[c] - (d + cfs) * q[c];
  }
}

/* point_in_pml + pml_damping_and_q (RHS_Interior.f90:2290-2451) */
static int point_in_pml(const orc_block *b, int x, int y, int z) {
  for (int f = 0; f < 6; ++f)
    if (b->pml[f] && pml_slab_contains(b, f, x, y, z)) return 1;
  return 0;
}

static void pml_damping_and_q(const orc_block *b, int x, int y, int z, double d[3], double Qd[3][9]) {
  long N = (long)b->nx * b->ny * b->nz;
  long id = IDX(b, x, y, z);
  memset(Qd, 0, sizeof(double) * 27);
  d[0] = d[1] = d[2] = 0.0;
  for (int f = 0; f < 6; ++f) {
    if (!b->pml[f] || !pml_slab_contains(b, f, x, y, z)) continue;
    int a = f / 2;
    d[a] = pml_d(b, f, x, y, z);
    for (int c = 0; c < 9; ++c) Qd[a][c] = b->pQ[f][c * N + id];
  }
}

/* apply_anelastic_Q{,8}_point_dispatch (RHS_Interior.f90:2599-2906) */
static void anelastic_point(orc_block *b, int x, int y, int z, const double Ux[9],
                            const double Uy[9], const double Uz[9], double DU[9]) {
  long id = IDX(b, x, y, z);
  for (int c6 = 0; c6 < 6; ++c6) DU[3 + c6] = DU[3 + c6] - eta_sum(b, c6, id);
  if (point_in_pml(b, x, y, z)) {
    double d[3], Qd[3][9];
    pml_damping_and_q(b, x, y, z, d, Qd);
    double exx = Ux[0] - d[0] * Qd[0][3];
    double eyy = Uy[1] - d[1] * Qd[1][4];
    double ezz = Uz[2] - d[2] * Qd[2][5];
    double tr = exx + eyy + ezz;
    double gxy = (Uy[0] + Ux[1]) - d[0] * Qd[0][6] - d[1] * Qd[1][6];
    double gxz = (Uz[0] + Ux[2]) - d[0] * Qd[0][7] - d[2] * Qd[2][7];
    double gyz = (Uz[1] + Uy[2]) - d[1] * Qd[1][8] - d[2] * Qd[2][8];
    q_rates(b, id, exx, eyy, ezz, tr, gxy, gxz, gyz);
  } else {
    double tr = Ux[0] + Uy[1] + Uz[2];
    q_rates(b, id, Ux[0], Uy[1], Uz[2], tr, Uy[0] + Ux[1], Uz[0] + Ux[2], Uz[1] + Uy[2]);
  }
}

/* RHS_near_boundaries, traditional branch (RHS_Interior.f90:136-281) */
static void rhs_near_boundaries(orc_block *b) {
  long N = (long)b->nx * b->ny * b->nz;
  int ix, fx, iy, fy, iz, fz;
  interior_extents(b, &ix, &fx, &iy, &fy, &iz, &fz);
#pragma omp parallel for collapse(2) schedule(static)
  for (int z = 1; z <= b->nz; ++z)
    for (int y = 1; y <= b->ny; ++y)
      for (int x = 1; x <= b->nx; ++x) {
        if ((iz + 1 <= z && z <= b->nz - fz) && (iy + 1 <= y && y <= b->ny - fy) &&
            (ix + 1 <= x && x <= b->nx - fx))
          continue;
        double Ux[9], Uy[9], Uz[9], DU[9];
        jju_x6(b, x, y, z, b->metx, Ux);
        jju_x6(b, x, y, z, b->mety, Uy);
        jju_x6(b, x, y, z, b->metz, Uz);
        long id = IDX(b, x, y, z);
        double lam = b->M[id], mu = b->M[N + id], rho = b->M[2 * N + id];
        double rhoJ_inv = 1.0 / (rho * b->J[id]);
        double lambda2mu = lam + 2.0 * mu;
        DU[0] = (Ux[3] + Uy[6] + Uz[7]) * rhoJ_inv;
        DU[1] = (Ux[6] + Uy[4] + Uz[8]) * rhoJ_inv;
        DU[2] = (Ux[7] + Uy[8] + Uz[5]) * rhoJ_inv;
        DU[3] = lambda2mu * Ux[0] + lam * (Uy[1] + Uz[2]);
        DU[4] = lambda2mu * Uy[1] + lam * (Ux[0] + Uz[2]);
        DU[5] = lambda2mu * Uz[2] + lam * (Ux[0] + Uy[1]);
        DU[6] = mu * (Uy[0] + Ux[1]);
        DU[7] = mu * (Uz[0] + Ux[2]);
        DU[8] = mu * (Uz[1] + Uy[2]);
        if (b->nmech) anelastic_point(b, x, y, z, Ux, Uy, Uz, DU);
        pml_point(b, DU, Ux, Uy, Uz, x, y, z, rhoJ_inv, lambda2mu);
        for (int c = 0; c < 9; ++c) b->DF[c * N + id] = b->DF[c * N + id] + DU[c];
      }
}

void orc_set_rates(orc_block *b) {
  rhs_center(b);
  rhs_near_boundaries(b);
}

/* ---------------------------------------------------------------------------
 * moment-tensor source (moment_tensor.f90:813-1124)
 * ------------------------------------------------------------------------- */
double orc_source_time(int type, double t, double t0, double w) {
  const double pi = 3.141592653589793;
  switch (type) {
    case ORC_SRC_GAUSSIAN:
      return (w / (sqrt(2.0 * pi))) * exp((-(w * w) * ((t - t0) * (t - t0))) / 2.0);
    case ORC_SRC_COSBELL:
      return (t < t0) ? (1.0 - cos(2.0 * pi * t / t0)) / t0 : 0.0;
    case ORC_SRC_RICKER: {
      /* Ricker wavelet (not in the reference): w = peak frequency f0, t0 = delay */
      double a = pi * w * (t - t0);
      return (1.0 - 2.0 * a * a) * exp(-a * a);
    }
  }
  return 0.0;
}

static double singular_source(int i, int i0, double h, double a, int order, int alpha) {
  if (order == 2) {
    double beta = (1.0 - 2.0 * a) / 4.0 + 0.25;
    double gam = (2.0 * a - 1.0) / 16.0;
    double gap = (2.0 * a + 1.0) / 16.0;
    if (i == i0 - 2) return (-gam) / h;
    if (i == i0 - 1) return (beta + 3.0 * gam - gap) / h;
    if (i == i0) return (1.0 - a - 2.0 * beta - 3.0 * gam + 3.0 * gap) / h;
    if (i == i0 + 1) return (a + beta + gam - 3.0 * gap) / h;
    if (i == i0 + 2) return gap / h;
    return 0.0;
  }
  static const double W0[9] = {-0.01171875, -0.03125, 0.046875, 0.28125, 0.4296875,
                               0.28125, 0.046875, -0.03125, -0.01171875};
  static const double W25[9] = {-0.00246875, -0.02275, -0.014125, 0.16575, 0.4171875,
                                0.38675, 0.117875, -0.02975, -0.01846875};
  static const double W50[9] = {-0.00728125, -0.028, 0.014875, 0.224, 0.4265625,
                                0.336, 0.081375, -0.032, -0.01553125};
  static const double W100[9] = {-0.0095625, -0.02990625, 0.03053125, 0.25284375, 0.42890625,
                                 0.30903125, 0.06384375, -0.03196875, -0.01371875};
  const double *W = alpha == 25 ? W25 : alpha == 50 ? W50 : alpha == 100 ? W100 : W0;
  int k = i - i0 + 4;
  if (k < 0 || k > 8) return 0.0;
  return W[k] / h;
}

/* set_moment_tensor_smooth (moment_tensor.f90:1255-1330): a Gaussian of
 * standard deviation 2h per axis over every node, h the coordinate's
 * backward difference along its own index (forward at the first node), the
 * reference's truncated pi */
static void mollified_source(orc_block *b, const orc_source *S, double A) {
  const double pi = 3.14159265359;
  long N = (long)b->nx * b->ny * b->nz;
  const double *X = b->X;
  for (int n = 1; n <= b->nz; ++n)
    for (int m = 1; m <= b->ny; ++m)
      for (int l = 1; l <= b->nx; ++l) {
        long id = IDX(b, l, m, n);
        double hx = l == 1 ? X[IDX(b, 2, m, n)] - X[id] : X[id] - X[IDX(b, l - 1, m, n)];
        double hy = m == 1 ? X[N + IDX(b, l, 2, n)] - X[N + id]
                           : X[N + id] - X[N + IDX(b, l, m - 1, n)];
        double hz = n == 1 ? X[2 * N + IDX(b, l, m, 2)] - X[2 * N + id]
                           : X[2 * N + id] - X[2 * N + IDX(b, l, m, n - 1)];
        double sx = 2.0 * hx, sy = 2.0 * hy, sz = 2.0 * hz;
        double ux = (X[id] - S->loc[0]) / sx;
        double uy = (X[N + id] - S->loc[1]) / sy;
        double uz = (X[2 * N + id] - S->loc[2]) / sz;
        double dx = exp(-0.5 * (ux * ux)) / (sx * sqrt(2.0 * pi));
        double dy = exp(-0.5 * (uy * uy)) / (sy * sqrt(2.0 * pi));
        double dz = exp(-0.5 * (uz * uz)) / (sz * sqrt(2.0 * pi));
        double delta = dx * dy * dz;
        b->DF[3 * N + id] = b->DF[3 * N + id] - A * S->mXX * delta;
        b->DF[4 * N + id] = b->DF[4 * N + id] - A * S->mYY * delta;
        b->DF[5 * N + id] = b->DF[5 * N + id] - A * S->mZZ * delta;
        b->DF[6 * N + id] = b->DF[6 * N + id] - A * S->mXY * delta;
        b->DF[7 * N + id] = b->DF[7 * N + id] - A * S->mXZ * delta;
        b->DF[8 * N + id] = b->DF[8 * N + id] - A * S->mYZ * delta;
      }
}

void orc_set_moment_tensor(orc_block *b, double t) {
  long N = (long)b->nx * b->ny * b->nz;
  for (int s = 0; s < b->nsrc; ++s) {
    const orc_source *S = &b->src[s];
    double A = orc_source_time(S->type, t, S->t_init, S->duration);
    if (S->order == 0) {
      mollified_source(b, S, A);
      continue;
    }
    int r = S->order == 2 ? 2 : 4;
    for (int n = S->near[2] - r; n <= S->near[2] + r; ++n)
      for (int m = S->near[1] - r; m <= S->near[1] + r; ++m)
        for (int l = S->near[0] - r; l <= S->near[0] + r; ++l) {
          if (l < 1 || l > b->nx || m < 1 || m > b->ny || n < 1 || n > b->nz) continue;
          long id = IDX(b, l, m, n);
          double hx = b->hq / b->metx[0 * N + id];
          double hy = b->hr / b->mety[1 * N + id];
          double hz = b->hs / b->metz[2 * N + id];
          double ax = (S->loc[0] - S->near_phys[0]) / hx;
          double ay = (S->loc[1] - S->near_phys[1]) / hy;
          double az = (S->loc[2] - S->near_phys[2]) / hz;
          double dx = singular_source(l, S->near[0], hx, ax, S->order, S->alpha);
          double dy = singular_source(m, S->near[1], hy, ay, S->order, S->alpha);
          double dz = singular_source(n, S->near[2], hz, az, S->order, S->alpha);
          double delta = dx * dy * dz;
          b->DF[3 * N + id] = b->DF[3 * N + id] - A * S->mXX * delta;
          b->DF[4 * N + id] = b->DF[4 * N + id] - A * S->mYY * delta;
          b->DF[5 * N + id] = b->DF[5 * N + id] - A * S->mZZ * delta;
          b->DF[6 * N + id] = b->DF[6 * N + id] - A * S->mXY * delta;
          b->DF[7 * N + id] = b->DF[7 * N + id] - A * S->mXZ * delta;
          b->DF[8 * N + id] = b->DF[8 * N + id] - A * S->mYZ * delta;
        }
  }
}

/* ---------------------------------------------------------------------------
 * SAT boundary terms: BC_* / BCm / BCp (BoundaryConditions.f90:50-924),
 * structure (RHS_Interior.f90:2267-2287),
 * impose_boundary_condition (RHS_Interior.f90:1584-1768)
 * ------------------------------------------------------------------------- */
static void bc_char(double ubc[9], const double u[9], const double *l, const double *m,
                    const double *n, double norm, double Zs, double Zp, double r, int plus) {
  double T[3], v[3] = {u[0], u[1], u[2]};
  T[0] = n[0] * u[3] + n[1] * u[6] + n[2] * u[7];
  T[1] = n[0] * u[6] + n[1] * u[4] + n[2] * u[8];
  T[2] = n[0] * u[7] + n[1] * u[8] + n[2] * u[5];
  double vl = l[0] * v[0] + l[1] * v[1] + l[2] * v[2];
  double vm = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  double vn = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  double Tl = l[0] * T[0] + l[1] * T[1] + l[2] * T[2];
  double Tm = m[0] * T[0] + m[1] * T[1] + m[2] * T[2];
  double Tn = n[0] * T[0] + n[1] * T[1] + n[2] * T[2];
  double pl, pm, pn, ql, qm, qn;
  if (plus) {
    pl = 0.5 * ((1.0 - r) * Zs * vl + (1.0 + r) * Tl);
    pm = 0.5 * ((1.0 - r) * Zs * vm + (1.0 + r) * Tm);
    pn = 0.5 * ((1.0 - r) * Zp * vn + (1.0 + r) * Tn);
    ql = 0.5 * ((1.0 - r) * vl + (1.0 + r) / Zs * Tl);
    qm = 0.5 * ((1.0 - r) * vm + (1.0 + r) / Zs * Tm);
    qn = 0.5 * ((1.0 - r) * vn + (1.0 + r) / Zp * Tn);
  } else {
    pl = 0.5 * ((1.0 - r) * Zs * vl - (1.0 + r) * Tl);
    pm = 0.5 * ((1.0 - r) * Zs * vm - (1.0 + r) * Tm);
    pn = 0.5 * ((1.0 - r) * Zp * vn - (1.0 + r) * Tn);
    ql = 0.5 * ((1.0 - r) * vl - (1.0 + r) / Zs * Tl);
    qm = 0.5 * ((1.0 - r) * vm - (1.0 + r) / Zs * Tm);
    qn = 0.5 * ((1.0 - r) * vn - (1.0 + r) / Zp * Tn);
  }
  double Fx = l[0] * pl + m[0] * pm + n[0] * pn;
  double Fy = l[1] * pl + m[1] * pm + n[1] * pn;
  double Fz = l[2] * pl + m[2] * pm + n[2] * pn;
  double F_x = l[0] * ql + m[0] * qm + n[0] * qn;
  double F_y = l[1] * ql + m[1] * qm + n[1] * qn;
  double F_z = l[2] * ql + m[2] * qm + n[2] * qn;
  double sg = plus ? 1.0 : -1.0;
  ubc[0] = norm * Fx;
  ubc[1] = norm * Fy;
  ubc[2] = norm * Fz;
  ubc[3] = sg * norm * n[0] * F_x;
  ubc[4] = sg * norm * n[1] * F_y;
  ubc[5] = sg * norm * n[2] * F_z;
  ubc[6] = sg * norm * (n[1] * F_x + n[0] * F_y);
  ubc[7] = sg * norm * (n[2] * F_x + n[0] * F_z);
  ubc[8] = sg * norm * (n[2] * F_y + n[1] * F_z);
}

/* the alpha/theta form of the penalty (commented out at
 * BoundaryConditions.f90:833-841 for BCp and 905-913 for BCm): the rate
 * vector itself, without structure(); alpha, theta per code as set in
 * BC_Lx..BC_Rz (BoundaryConditions.f90:129-160) */
static void bc_alpha_theta(double Ux[9], const double *l, const double *m, const double *n,
                           const double u[9], double norm, double Zs, double Zp, double r,
                           int code, int plus, double lam, double mu, double rho) {
  const double alpha = code == 1 ? 2.0 : (code == 2 ? 1.0 : 0.0);
  const double theta = code == 3 ? -1.0 : 0.0;
  double T[3], v[3] = {u[0], u[1], u[2]};
  T[0] = n[0] * u[3] + n[1] * u[6] + n[2] * u[7];
  T[1] = n[0] * u[6] + n[1] * u[4] + n[2] * u[8];
  T[2] = n[0] * u[7] + n[1] * u[8] + n[2] * u[5];
  double vl = l[0] * v[0] + l[1] * v[1] + l[2] * v[2];
  double vm = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  double vn = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
  double Tl = l[0] * T[0] + l[1] * T[1] + l[2] * T[2];
  double Tm = m[0] * T[0] + m[1] * T[1] + m[2] * T[2];
  double Tn = n[0] * T[0] + n[1] * T[1] + n[2] * T[2];
  const double sg = plus ? 1.0 : -1.0;
  double pl = 0.5 * ((1.0 - r) * Zs * vl + sg * (1.0 + r) * Tl);
  double pm = 0.5 * ((1.0 - r) * Zs * vm + sg * (1.0 + r) * Tm);
  double pn = 0.5 * ((1.0 - r) * Zp * vn + sg * (1.0 + r) * Tn);
  double ql = 0.5 * ((1.0 - r) * vl + sg * (1.0 + r) / Zs * Tl);
  double qm = 0.5 * ((1.0 - r) * vm + sg * (1.0 + r) / Zs * Tm);
  double qn = 0.5 * ((1.0 - r) * vn + sg * (1.0 + r) / Zp * Tn);
  double Fx = l[0] * pl + m[0] * pm + n[0] * pn;
  double Fy = l[1] * pl + m[1] * pm + n[1] * pn;
  double Fz = l[2] * pl + m[2] * pm + n[2] * pn;
  double F_x = l[0] * ql + m[0] * qm + n[0] * qn;
  double F_y = l[1] * ql + m[1] * qm + n[1] * qn;
  double F_z = l[2] * ql + m[2] * qm + n[2] * qn;
  double l2m = lam + 2.0 * mu;
  Ux[0] = alpha * norm / rho * Fx;
  Ux[1] = alpha * norm / rho * Fy;
  Ux[2] = alpha * norm / rho * Fz;
  Ux[3] = theta * norm * (l2m * n[0] * F_x + lam * (n[1] * F_y + n[2] * F_z));
  Ux[4] = theta * norm * (l2m * n[1] * F_y + lam * (n[0] * F_x + n[2] * F_z));
  Ux[5] = theta * norm * (l2m * n[2] * F_z + lam * (n[0] * F_x + n[1] * F_y));
  Ux[6] = theta * mu * norm * (n[1] * F_x + n[0] * F_y);
  Ux[7] = theta * mu * norm * (n[2] * F_x + n[0] * F_z);
  Ux[8] = theta * mu * norm * (n[2] * F_y + n[1] * F_z);
}

void orc_impose_boundary_condition(orc_block *b) {
  long N = (long)b->nx * b->ny * b->nz;
  double h[3] = {b->hq, b->hr, b->hs};
  for (int f = 0; f < 6; ++f) {
    int code = b->bc[f];
    if (code <= 0) continue;
    double r = code == 1 ? 0.0 : (code == 2 ? 1.0 : -1.0);
    int axis = f / 2, plus = f % 2;
    int fixed = plus ? (axis == 0 ? b->nx : (axis == 1 ? b->ny : b->nz)) : 1;
    int n1 = axis == 0 ? b->ny : b->nx;
    int n2 = axis == 2 ? b->ny : b->nz;
    long NF = (long)n1 * n2;
    for (int j = 1; j <= n2; ++j)
      for (int i = 1; i <= n1; ++i) {
        int x, y, z;
        if (axis == 0) { x = fixed; y = i; z = j; }
        else if (axis == 1) { x = i; y = fixed; z = j; }
        else { x = i; y = j; z = fixed; }
        long id = IDX(b, x, y, z);
        long fid = (long)(i - 1) + (long)n1 * (j - 1);
        double lam = b->M[id], mu = b->M[N + id], rho = b->M[2 * N + id];
        double cp = sqrt((2.0 * mu + lam) / rho), cs = sqrt(mu / rho);
        double Zp = rho * cp, Zs = rho * cs;
        double qx = b->metx[axis * N + id], qy = b->mety[axis * N + id], qz = b->metz[axis * N + id];
        double norm = sqrt(qx * qx + qy * qy + qz * qz);
        double l[3], m[3], n[3], u[9], ubc[9], Ux[9];
        for (int k = 0; k < 3; ++k) {
          l[k] = b->nl[f][k * NF + fid];
          m[k] = b->nm[f][k * NF + fid];
          n[k] = b->nn[f][k * NF + fid];
        }
        for (int c = 0; c < 9; ++c) u[c] = b->F[c * N + id];
        bc_char(ubc, u, l, m, n, norm, Zs, Zp, r, plus);
        Ux[0] = 1.0 / rho * ubc[0];
        Ux[1] = 1.0 / rho * ubc[1];
        Ux[2] = 1.0 / rho * ubc[2];
        Ux[3] = (2.0 * mu + lam) * ubc[3] + lam * (ubc[4] + ubc[5]);
        Ux[4] = (2.0 * mu + lam) * ubc[4] + lam * (ubc[3] + ubc[5]);
        Ux[5] = (2.0 * mu + lam) * ubc[5] + lam * (ubc[3] + ubc[4]);
        Ux[6] = mu * ubc[6];
        Ux[7] = mu * ubc[7];
        Ux[8] = mu * ubc[8];
        if (b->sat_form == 1) bc_alpha_theta(Ux, l, m, n, u, norm, Zs, Zp, r, code, plus, lam, mu, rho);
        double pen = b->tau0 / h[axis];
        for (int c = 0; c < 9; ++c) b->DF[c * N + id] = b->DF[c * N + id] - pen * Ux[c];
        if (b->pml[f])
          for (int c = 0; c < 9; ++c) b->pDQ[f][c * N + id] = b->pDQ[f][c * N + id] - pen * ubc[c];
      }
  }
}

/* ---------------------------------------------------------------------------
 * time_step_RK (time_step.f90:68-150)
 * ------------------------------------------------------------------------- */
void orc_time_step(orc_block *b, double dt, int nsteps, double *t,
                   int nrec, const long *rec_idx, double *seis) {
  double A[5], B[5], C[6];
  long N = (long)b->nx * b->ny * b->nz;
  orc_rk_coefficients(A, B, C);
  for (int n = 0; n < nsteps; ++n) {
    double t0 = *t;
    for (int s = 0; s < 5; ++s) {
      orc_scale_rates(b, A[s]);
      orc_set_rates(b);
      if (b->nsrc) orc_set_moment_tensor(b, *t);
      orc_impose_boundary_condition(b);
      if (s == 0 && nrec > 0)
        for (int r = 0; r < nrec; ++r)
          for (int c = 0; c < 9; ++c)
            seis[((long)n * nrec + r) * 9 + c] = b->F[c * N + rec_idx[r]];
      *t = t0 + C[s + 1] * dt;
      orc_update_fields(b, B[s] * dt);
    }
  }
}
