/*
 * orc_mm10.c -- oracle: mm10 crystal plasticity (Voce hardening, FCC or BCC
 * 12-system slip) for one Gauss point, with its consistent tangent.
 *
 * TEST INFRASTRUCTURE ONLY (see w3d_oracle.h).  Restates, in the reference's
 * own operation order:
 *   mod_crystals.f:238-417 (initialize_new_crystal defaults), :421-500, :560
 *     (fcc / bcc slip systems), :1795-1840 (elastic flexibility, isotropic and
 *     cubic, inverted by mm10_invsym)
 *   rknstr.f:1568-1640 (setup_mm10_rknstr: Schmid tensors ms/qs of each slip
 *     system and the stiffness rotated into the reference frame)
 *   mm10_a.f:29 (mm10: loop over crystals, averages, mm10_a_store_crystal),
 *     :204 (mm10_a_do_crystal), :480 (mm10_init_cc_hist0), :1777, :1808,
 *     :2605 (history initialisation at step 1), :2493 (mm10_copy_cc_hist),
 *     :1689 (mm10_setup_np1), :852 (mm10_setup), :2071, :2086 (Voce init
 *     and setup), :1102 (mm10_solve_crystal), :2657 (mm10_solve_strup with
 *     sub-stepping), :2901 (mm10_solve: predictor :3032, corrector :3224),
 *     :664 (mm10_tangent), :3991 (mm10_a_make_symm_1), :3375
 *     (mm10_update_rotation), :3410 (mm10_wv2wt), :3447 (mm10_expmw3x3),
 *     :3493 (mm10_output), :1195 (mm10_update_euler_angles), :998
 *     (mm10_store_cryhist), :1310 (mm10_rotation_matrix), :1431
 *     (mm10_rt2rve), :1492 (mm10_rt2rvw), :1523 (mm10_et2ev), :1553
 *     (mm10_wt2wv)
 *   mm10_b.f:63 (formR2), :177 (formJ11), :281 (formJ12), :377 (formJ21),
 *     :439 (formJ22), :909 (formJ), :1037 (formR), :1073 (formR1), :1169
 *     (form_dbarp), :1239 (form_wbarp), :1312 (form_wp), :1513 (symsw),
 *     :1576 (symswmat), :1613 (iw), :1813 (slipinc), :1845 (rs), :1896-2074
 *     (Voce: h, estress, ehard, ed, dgdt, dgdh, dgdd)
 *   rstgp1.f:4327 (drive_10_update: small-strain rotation = I, the
 *     "extrapolate off" flag of iteration 0)
 *
 * LAPACK/BLAS calls of the reference (DGESV, DGER, DGEMM, DPOSV, DSYTRF/
 * DSYTRI) are restated with the unblocked reference-BLAS loop orders:
 * partial-pivot LU (first largest pivot), forward/back substitution, an
 * upper Cholesky for DPOSV and an LU inverse in place of the symmetric
 * Bunch-Kaufman inverse.  Their results agree with any LAPACK to rounding.
 *
 * Scope: hardening model 1 (Voce) with k_0 = 0 (no geometric hardening),
 * small-displacement form (rotation R = I), no temperatures, no nonlocal
 * state, symmetric assembly (the tangent is symmetrised).
 *
 * History of one point (0-based word offsets; 87 + 71 words per crystal):
 *   0..35 tangent (6x6, column-major), 36..62 grad Fe (never written by the
 *   reference for small strain; written as zeros here), 63..71 R, 72..74
 *   work / plastic work / plastic strain sums, 75..86 slip totals;
 *   crystal c at 87 + 71c: +0 stress(6), +6 Euler angles(3), +9 Rp(9,
 *   column-major), +18 D(6), +24 elastic strain(6), +30 slip increments(12),
 *   +42 tau_tilde, +43 user words(15), +58 tau_tilde rate, +59 ep(6),
 *   +65 ed(6).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "w3d_oracle.h"

#define NSL 12
#define NCOMMON 87
#ifndef MXGP
#define MXGP 27
#endif
#define NCRYS 71

/* mm10_constants (mod_crystals.f:122-129) */
static const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0,
                    half = 0.5, pi = 3.141592653589793, one_eighty = 180.0,
                    onept5 = 1.5, pt1 = 0.1;

/* slip systems, mod_crystals.f:441-500: fcc b in units of 1/sqrt(2), n in
   units of 1/sqrt(3); bcc swaps the two tables (mod_crystals.f:502-600) */
static const int FCC_B[NSL][3] = {{0, -1, 1}, {1, 0, -1}, {-1, 1, 0},
                                  {0, 1, 1},  {1, 0, 1},  {1, -1, 0},
                                  {0, -1, 1}, {-1, 0, -1}, {1, 1, 0},
                                  {0, 1, 1},  {1, 0, -1}, {-1, -1, 0}};
static const int FCC_N[NSL][3] = {{1, 1, 1},  {1, 1, 1},   {1, 1, 1},
                                  {-1, -1, 1}, {-1, -1, 1}, {-1, -1, 1},
                                  {-1, 1, 1}, {-1, 1, 1},   {-1, 1, 1},
                                  {1, -1, 1}, {1, -1, 1},   {1, -1, 1}};

/* ------------------------------------------------------------------ */
/* small dense linear algebra (column-major, a(i,j) at a[i + n*j])      */
/* ------------------------------------------------------------------ */
/* DGESV: LU with partial pivoting (dgetf2 order), then dgetrs. */
/* Work counters of the local solves (test and measurement instrument):
   [0] crystal updates, [1] sub-step solves, [2] formR1 evaluations (also
   inside formR), [3] formR, [4] formJ11 (also inside formJ), [5] formJ,
   [6] 6x6 LU solves, [7] 7x7 LU solves; and a histogram of formR1
   evaluations per point (bin 63 = 63 or more). */
static long long g_cnt[8];
static long long g_hist[64];
static __thread long long t_evals;
#define CNT(i) __atomic_fetch_add(&g_cnt[i], 1, __ATOMIC_RELAXED)
void orc_mm10_counters(long long* out8, long long* hist64, int reset) {
  for (int i = 0; i < 8; ++i) {
    if (out8) out8[i] = __atomic_load_n(&g_cnt[i], __ATOMIC_RELAXED);
    if (reset) __atomic_store_n(&g_cnt[i], 0, __ATOMIC_RELAXED);
  }
  for (int i = 0; i < 64; ++i) {
    if (hist64) hist64[i] = __atomic_load_n(&g_hist[i], __ATOMIC_RELAXED);
    if (reset) __atomic_store_n(&g_hist[i], 0, __ATOMIC_RELAXED);
  }
}

static int lu_solve(int n, double* a, int nrhs, double* b) {
  if (n == 6 && nrhs == 1) CNT(6);
  if (n == 7) CNT(7);
  int ipiv[8];
  for (int j = 0; j < n; ++j) {
    int p = j;
    double amax = fabs(a[j + n * j]);
    for (int i = j + 1; i < n; ++i)
      if (fabs(a[i + n * j]) > amax) {
        amax = fabs(a[i + n * j]);
        p = i;
      }
    ipiv[j] = p;
    if (a[p + n * j] == 0.0) return -1;
    if (p != j)
      for (int k = 0; k < n; ++k) {
        double t = a[j + n * k];
        a[j + n * k] = a[p + n * k];
        a[p + n * k] = t;
      }
    double r = one / a[j + n * j];
    for (int i = j + 1; i < n; ++i) a[i + n * j] = a[i + n * j] * r;
    for (int k = j + 1; k < n; ++k) {
      double t = -a[j + n * k];
      if (t != 0.0)
        for (int i = j + 1; i < n; ++i)
          a[i + n * k] = a[i + n * k] + a[i + n * j] * t;
    }
  }
  for (int c = 0; c < nrhs; ++c) {
    double* x = b + n * c;
    for (int j = 0; j < n; ++j)
      if (ipiv[j] != j) {
        double t = x[j];
        x[j] = x[ipiv[j]];
        x[ipiv[j]] = t;
      }
    for (int k = 0; k < n; ++k)
      if (x[k] != 0.0)
        for (int i = k + 1; i < n; ++i) x[i] = x[i] - x[k] * a[i + n * k];
    for (int k = n - 1; k >= 0; --k)
      if (x[k] != 0.0) {
        x[k] = x[k] / a[k + n * k];
        for (int i = 0; i < k; ++i) x[i] = x[i] - x[k] * a[i + n * k];
      }
  }
  return 0;
}

/* DPOSV('U', 6, 1): Cholesky A = U^T U (dpotf2 order), two triangular
   solves. */
static int chol_solve6(const double* a_in, double* b) {
  double a[36];
  memcpy(a, a_in, sizeof a);
  for (int j = 0; j < 6; ++j) {
    double s = 0.0;
    for (int k = 0; k < j; ++k) s = s + a[k + 6 * j] * a[k + 6 * j];
    double ajj = a[j + 6 * j] - s;
    if (ajj <= 0.0) return -1;
    ajj = sqrt(ajj);
    a[j + 6 * j] = ajj;
    for (int c = j + 1; c < 6; ++c) {
      double t = 0.0;
      for (int k = 0; k < j; ++k) t = t + a[k + 6 * j] * a[k + 6 * c];
      a[j + 6 * c] = (a[j + 6 * c] - t) / ajj;
    }
  }
  for (int j = 0; j < 6; ++j) { /* U^T y = b */
    double t = b[j];
    for (int k = 0; k < j; ++k) t = t - a[k + 6 * j] * b[k];
    b[j] = t / a[j + 6 * j];
  }
  for (int j = 5; j >= 0; --j) { /* U x = y */
    b[j] = b[j] / a[j + 6 * j];
    for (int i = 0; i < j; ++i) b[i] = b[i] - b[j] * a[i + 6 * j];
  }
  return 0;
}

/* mm10_invsym: inverse of a symmetric 6x6, full storage on return */
static int inv6(double* a) {
  double lu[36], id[36];
  memcpy(lu, a, sizeof lu);
  memset(id, 0, sizeof id);
  for (int i = 0; i < 6; ++i) id[i + 6 * i] = one;
  if (lu_solve(6, lu, 6, id)) return -1;
  memcpy(a, id, sizeof id);
  return 0;
}

/* ------------------------------------------------------------------ */
/* tensor <-> vector conversions and rotations                          */
/* ------------------------------------------------------------------ */
#define M3(a, i, j) (a)[(i) + 3 * (j)]
#define M6(a, i, j) (a)[(i) + 6 * (j)]

/* mm10_rotation_matrix, Kocks convention (mm10_a.f:1310) */
void orc_mm10_rotation_matrix(const double angles[3], int angle_type,
                              double r[9]) {
  double a = angles[0], b = angles[1], c = angles[2];
  if (angle_type == 1) { /* degrees */
    a = a * pi / one_eighty;
    b = b * pi / one_eighty;
    c = c * pi / one_eighty;
  }
  double psi = a, theta = b, phi = c;
  M3(r, 0, 0) = -sin(psi) * sin(phi) - cos(psi) * cos(phi) * cos(theta);
  M3(r, 0, 1) = cos(psi) * sin(phi) - sin(psi) * cos(phi) * cos(theta);
  M3(r, 0, 2) = cos(phi) * sin(theta);
  M3(r, 1, 0) = sin(psi) * cos(phi) - cos(psi) * sin(phi) * cos(theta);
  M3(r, 1, 1) = -cos(psi) * cos(phi) - sin(psi) * sin(phi) * cos(theta);
  M3(r, 1, 2) = sin(phi) * sin(theta);
  M3(r, 2, 0) = cos(psi) * sin(theta);
  M3(r, 2, 1) = sin(psi) * sin(theta);
  M3(r, 2, 2) = cos(theta);
}

/* mm10_rt2rve (mm10_a.f:1431) */
void orc_mm10_rt2rve(const double* rt, double* rv) {
#define RT(i, j) M3(rt, (i)-1, (j)-1)
#define RV(i, j) M6(rv, (i)-1, (j)-1)
  RV(1, 1) = RT(1, 1) * RT(1, 1);
  RV(1, 2) = RT(1, 2) * RT(1, 2);
  RV(1, 3) = RT(1, 3) * RT(1, 3);
  RV(1, 4) = two * RT(1, 1) * RT(1, 2);
  RV(1, 5) = two * RT(1, 3) * RT(1, 2);
  RV(1, 6) = two * RT(1, 1) * RT(1, 3);
  RV(2, 1) = RT(2, 1) * RT(2, 1);
  RV(2, 2) = RT(2, 2) * RT(2, 2);
  RV(2, 3) = RT(2, 3) * RT(2, 3);
  RV(2, 4) = two * RT(2, 1) * RT(2, 2);
  RV(2, 5) = two * RT(2, 3) * RT(2, 2);
  RV(2, 6) = two * RT(2, 1) * RT(2, 3);
  RV(3, 1) = RT(3, 1) * RT(3, 1);
  RV(3, 2) = RT(3, 2) * RT(3, 2);
  RV(3, 3) = RT(3, 3) * RT(3, 3);
  RV(3, 4) = two * RT(3, 1) * RT(3, 2);
  RV(3, 5) = two * RT(3, 3) * RT(3, 2);
  RV(3, 6) = two * RT(3, 1) * RT(3, 3);
  RV(4, 1) = RT(1, 1) * RT(2, 1);
  RV(4, 2) = RT(1, 2) * RT(2, 2);
  RV(4, 3) = RT(1, 3) * RT(2, 3);
  RV(4, 4) = RT(1, 1) * RT(2, 2) + RT(2, 1) * RT(1, 2);
  RV(4, 5) = RT(1, 2) * RT(2, 3) + RT(1, 3) * RT(2, 2);
  RV(4, 6) = RT(1, 1) * RT(2, 3) + RT(1, 3) * RT(2, 1);
  RV(5, 1) = RT(2, 1) * RT(3, 1);
  RV(5, 2) = RT(3, 2) * RT(2, 2);
  RV(5, 3) = RT(2, 3) * RT(3, 3);
  RV(5, 4) = RT(2, 1) * RT(3, 2) + RT(2, 2) * RT(3, 1);
  RV(5, 5) = RT(2, 2) * RT(3, 3) + RT(3, 2) * RT(2, 3);
  RV(5, 6) = RT(2, 1) * RT(3, 3) + RT(2, 3) * RT(3, 1);
  RV(6, 1) = RT(1, 1) * RT(3, 1);
  RV(6, 2) = RT(1, 2) * RT(3, 2);
  RV(6, 3) = RT(1, 3) * RT(3, 3);
  RV(6, 4) = RT(1, 1) * RT(3, 2) + RT(1, 2) * RT(3, 1);
  RV(6, 5) = RT(1, 2) * RT(3, 3) + RT(1, 3) * RT(3, 2);
  RV(6, 6) = RT(1, 1) * RT(3, 3) + RT(3, 1) * RT(1, 3);
#undef RV
}

/* mm10_rt2rvw (mm10_a.f:1492) */
void orc_mm10_rt2rvw(const double* rt, double* rv) {
#define RW(i, j) M3(rv, (i)-1, (j)-1)
  RW(1, 1) = RT(2, 2) * RT(3, 3) - RT(2, 3) * RT(3, 2);
  RW(1, 2) = RT(2, 1) * RT(3, 3) - RT(2, 3) * RT(3, 1);
  RW(1, 3) = RT(2, 1) * RT(3, 2) - RT(2, 2) * RT(3, 1);
  RW(2, 1) = RT(1, 2) * RT(3, 3) - RT(1, 3) * RT(3, 2);
  RW(2, 2) = RT(1, 1) * RT(3, 3) - RT(1, 3) * RT(3, 1);
  RW(2, 3) = RT(1, 1) * RT(3, 2) - RT(1, 2) * RT(3, 1);
  RW(3, 1) = RT(1, 2) * RT(2, 3) - RT(1, 3) * RT(2, 2);
  RW(3, 2) = RT(1, 1) * RT(2, 3) - RT(1, 3) * RT(2, 1);
  RW(3, 3) = RT(1, 1) * RT(2, 2) - RT(1, 2) * RT(2, 1);
#undef RW
#undef RT
}

/* mm10_wv2wt (mm10_a.f:3410) and mm10_expmw3x3 (mm10_a.f:3447):
   exp of the skew tensor of wv by the Rodrigues formula */
void orc_mm10_expmw(const double wv[3], double a[9]) {
  double w[9];
  M3(w, 0, 0) = zero;   M3(w, 0, 1) = wv[2];  M3(w, 0, 2) = wv[1];
  M3(w, 1, 0) = -wv[2]; M3(w, 1, 1) = zero;   M3(w, 1, 2) = wv[0];
  M3(w, 2, 0) = -wv[1]; M3(w, 2, 1) = -wv[0]; M3(w, 2, 2) = zero;
  double alpha = sqrt(M3(w, 1, 2) * M3(w, 1, 2) + M3(w, 0, 2) * M3(w, 0, 2) +
                      M3(w, 0, 1) * M3(w, 0, 1));
  if (alpha < 1.0e-16) {
    for (int k = 0; k < 9; ++k) a[k] = zero;
  } else {
    /* DGEMM('N','N',3,3,3,(1-cos)/alpha^2, w, w, sin/alpha, a=w) */
    double ca = (one - cos(alpha)) / (alpha * alpha);
    double sa = sin(alpha) / alpha;
    for (int j = 0; j < 3; ++j) {
      for (int i = 0; i < 3; ++i) a[i + 3 * j] = sa * w[i + 3 * j];
      for (int l = 0; l < 3; ++l) {
        double t = ca * w[l + 3 * j];
        for (int i = 0; i < 3; ++i) a[i + 3 * j] = a[i + 3 * j] + t * w[i + 3 * l];
      }
    }
  }
  M3(a, 0, 0) = M3(a, 0, 0) + one;
  M3(a, 1, 1) = M3(a, 1, 1) + one;
  M3(a, 2, 2) = M3(a, 2, 2) + one;
}

/* [a] = [b][c], 3x3 (mm10_a_mult_type_1) */
static void mul33(double* a, const double* b, const double* c) {
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      a[i + 3 * j] = M3(b, i, 0) * M3(c, 0, j) + M3(b, i, 1) * M3(c, 1, j) +
                     M3(b, i, 2) * M3(c, 2, j);
}
/* a = [b] c, 6x6 (mm10_a_mult_type_2) */
static void mv6(double* a, const double* b, const double* c) {
  for (int i = 0; i < 6; ++i) a[i] = M6(b, i, 0) * c[0];
  for (int j = 1; j < 6; ++j)
    for (int i = 0; i < 6; ++i) a[i] = a[i] + M6(b, i, j) * c[j];
}
/* a = [b] c + k d (mm10_a_mult_type_4 / mm10_b_mult_type_4) */
static void mv6_plus(double* a, const double* b, const double* c,
                     const double* d, double k) {
  for (int i = 0; i < 6; ++i) a[i] = k * d[i];
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) a[i] = a[i] + M6(b, i, j) * c[j];
}
/* a = [b] c, 3x3 (mm10_a_mult_type_3) */
static void mv3(double* a, const double* b, const double* c) {
  for (int i = 0; i < 3; ++i)
    a[i] = M3(b, i, 0) * c[0] + M3(b, i, 1) * c[1] + M3(b, i, 2) * c[2];
}

/* ------------------------------------------------------------------ */
/* crystal set-up                                                       */
/* ------------------------------------------------------------------ */
int orc_mm10_crystal(const w3d_material* mat, const double angles[3],
                     orc_crystal* cr) {
  if (mat->slip_type != 1 && mat->slip_type != 2) return -1;
  if (mat->elastic_type != 1 && mat->elastic_type != 2) return -1;
  /* finalize_new_crystal: flexibility, inverted (mod_crystals.f:1795) */
  double e = mat->cp_e, v = mat->cp_nu, u = mat->cp_mu;
  double fl[36];
  memset(fl, 0, sizeof fl);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M6(fl, i, j) = i == j ? 1 / e : -v / e;
  double sh = mat->elastic_type == 1 ? 2 * (1 + v) / e : 1 / u;
  for (int i = 3; i < 6; ++i) M6(fl, i, i) = sh;
  double c0[36];
  memcpy(c0, fl, sizeof c0);
  if (inv6(c0)) return -1;
  /* setup_mm10_rknstr_a(4): reference rotation; (5): slip tensors */
  orc_mm10_rotation_matrix(angles, mat->angle_type, cr->g);
  double tr[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M3(tr, i, j) = M3(cr->g, j, i);
  const double fb = 1 / sqrt(2.0), fn = 1 / sqrt(3.0);
  for (int s = 0; s < NSL; ++s) {
    double b[3], n[3], bs[3], ns[3];
    for (int k = 0; k < 3; ++k) {
      double bv = fb * FCC_B[s][k], nv = fn * FCC_N[s][k];
      if (FCC_B[s][k] == 0) bv = 0.0;
      if (mat->slip_type == 1) {
        b[k] = bv;
        n[k] = nv;
      } else { /* bcc: the fcc tables with b and n exchanged */
        b[k] = nv;
        n[k] = bv;
      }
    }
    mv3(bs, tr, b);
    mv3(ns, tr, n);
    double A[9], At[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        M3(A, i, j) = bs[i] * ns[j];
        M3(At, j, i) = M3(A, i, j);
      }
    double As[9], Aa[9];
    for (int k = 0; k < 9; ++k) {
      As[k] = 0.5 * (A[k] + At[k]);
      Aa[k] = 0.5 * (A[k] - At[k]);
    }
    /* mm10_et2ev, mm10_wt2wv */
    cr->ms[s][0] = M3(As, 0, 0);
    cr->ms[s][1] = M3(As, 1, 1);
    cr->ms[s][2] = M3(As, 2, 2);
    cr->ms[s][3] = two * M3(As, 0, 1);
    cr->ms[s][4] = two * M3(As, 1, 2);
    cr->ms[s][5] = two * M3(As, 0, 2);
    cr->qs[s][0] = M3(Aa, 1, 2);
    cr->qs[s][1] = M3(Aa, 0, 2);
    cr->qs[s][2] = M3(Aa, 0, 1);
  }
  /* stiffness rotated forward: C = Rstiff * (C0 * Rstiff^T) */
  double rs[36], rst[36], t66[36];
  orc_mm10_rt2rve(tr, rs);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) M6(rst, i, j) = M6(rs, j, i);
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) {
      double t = 0.0;
      for (int k = 0; k < 6; ++k) t = t + M6(c0, i, k) * M6(rst, k, j);
      M6(t66, i, j) = t;
    }
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) {
      double t = 0.0;
      for (int k = 0; k < 6; ++k) t = t + M6(rs, i, k) * M6(t66, k, j);
      M6(cr->C, i, j) = t;
    }
  memcpy(cr->angles, angles, sizeof cr->angles);
  return 0;
}

/* ------------------------------------------------------------------ */
/* crystal state (mm10_defs crystal_state) and material constants       */
/* ------------------------------------------------------------------ */
typedef struct {
  double R[9], Rp[9];
  double stress[6], D[6], eps[6], ed[6], ep[6];
  double euler[3];
  double slip_incs[NSL], tau_l[NSL];
  double tangent[36];
  double tau_tilde, tt_rate;
  double tinc, dg, mu_harden, work_inc, p_work_inc, p_strain_inc;
  double ms[NSL][6], qs[NSL][3], qc[NSL][3];
  double u[15];
} cstate;

typedef struct {
  const orc_crystal* cr;
  double rate_n, tau_y, tau_v, voce_m, theta_0, iD_v, eps_dot_0_y;
  double atol, atol1, rtol, rtol1;
  int miter, alter_mode, angle_type;
} cprops;

static double sgn(double x) { return x >= 0.0 ? one : -one; }


/* mm10_setup_np1 (mm10_a.f:1689); ed/ep keep their previous contents in the
   reference and are left untouched here too */
static void setup_np1(const double* Rur, const double* dstrain, double dt,
                      cstate* np1) {
  np1->tinc = dt;
  np1->dg = zero;
  np1->mu_harden = zero;
  np1->work_inc = np1->p_work_inc = np1->p_strain_inc = zero;
  for (int i = 0; i < 6; ++i) {
    np1->D[i] = dstrain[i];
    np1->stress[i] = zero;
    np1->eps[i] = zero;
  }
  np1->euler[0] = np1->euler[1] = np1->euler[2] = zero;
  for (int i = 0; i < NSL; ++i) np1->tau_l[i] = np1->slip_incs[i] = zero;
  np1->tau_tilde = np1->tt_rate = zero;
  for (int i = 0; i < 15; ++i) np1->u[i] = zero;
  for (int k = 0; k < 9; ++k) {
    np1->R[k] = Rur[k];
    np1->Rp[k] = zero;
  }
  memset(np1->tangent, 0, sizeof np1->tangent);
  memset(np1->ms, 0, sizeof np1->ms);
  memset(np1->qs, 0, sizeof np1->qs);
  memset(np1->qc, 0, sizeof np1->qc);
}

/* mm10_setup (mm10_a.f:852) + mm10_setup_voche (:2086); with k_0 = 0 the
   geometric-hardening tau_l are zero */
static void setup(const cprops* p, cstate* np1, const cstate* n) {
  const double twothirds = 2.0 / 3.0;
  double t1 = np1->D[0] * np1->D[0] + np1->D[1] * np1->D[1] +
              np1->D[2] * np1->D[2];
  double t2 = np1->D[3] * np1->D[3] + np1->D[4] * np1->D[4] +
              np1->D[5] * np1->D[5];
  np1->dg = sqrt(twothirds * (t1 + half * t2));
  double tn[9], RE[36], RW[9], RWC[9], work[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) M3(tn, i, j) = M3(n->Rp, j, i);
  orc_mm10_rt2rve(tn, RE);
  orc_mm10_rt2rvw(tn, RW);
  mul33(work, np1->R, tn);
  orc_mm10_rt2rvw(work, RWC);
  for (int i = 0; i < NSL; ++i) {
    mv6(np1->ms[i], RE, p->cr->ms[i]);
    mv3(np1->qs[i], RW, p->cr->qs[i]);
    mv3(np1->qc[i], RWC, p->cr->qs[i]);
  }
  np1->mu_harden = M6(p->cr->C, 5, 5);
  if (p->alter_mode) np1->dg = p->eps_dot_0_y * np1->tinc;
  for (int i = 0; i < NSL; ++i) np1->tau_l[i] = zero;
}

/* mm10_rs, mm10_slipinc (mm10_b.f:1845, :1813) */
static double rs_of(const cstate* np1, const double* s, int i) {
  const double* m = np1->ms[i];
  return s[0] * m[0] + s[1] * m[1] + s[2] * m[2] + s[3] * m[3] + s[4] * m[4] +
         s[5] * m[5];
}
static double slipinc(const cprops* p, const cstate* np1, const double* s,
                      double tt, int i) {
  double rs = rs_of(np1, s, i);
  return np1->dg / tt * pow(fabs(rs / tt), p->rate_n - one) * rs;
}

/* mm10_symsw (mm10_b.f:1513), kept as written */
static void symsw(const double* s, const double* w, double* sw) {
  sw[0] = s[3] * w[2] - s[5] * w[1];
  sw[1] = s[3] * w[2] - s[4] * w[0];
  sw[2] = s[5] * w[1] + s[4] * w[0];
  sw[3] = half * (w[2] * (s[0] - s[1]) + w[0] * s[5] - w[1] * s[4]);
  sw[4] = half * (w[0] * (s[1] - s[2]) + w[1] * s[3] + w[2] * s[5]);
  sw[5] = half * (w[1] * (s[0] - s[2]) + w[0] * s[3] - w[2] * s[4]);
}

/* mm10_form_dbarp / form_wbarp / form_wp for Voce */
static void form_dbarp(const cprops* p, const cstate* np1, const double* s,
                       double tt, double* dbar) {
  for (int k = 0; k < 6; ++k) dbar[k] = zero;
  for (int i = 0; i < NSL; ++i) {
    double gi = slipinc(p, np1, s, tt, i);
    double rs = rs_of(np1, s, i);
    double c = rs * np1->tinc * p->iD_v + gi;
    for (int k = 0; k < 6; ++k) dbar[k] = dbar[k] + c * np1->ms[i][k];
  }
}
static void form_w(const cprops* p, const cstate* np1, const double* s,
                   double tt, const double (*q)[3], double* w) {
  w[0] = w[1] = w[2] = zero;
  for (int i = 0; i < NSL; ++i) {
    double gi = slipinc(p, np1, s, tt, i);
    for (int k = 0; k < 3; ++k) w[k] = w[k] + gi * q[i][k];
    double rs = rs_of(np1, s, i);
    double c = rs * np1->tinc * p->iD_v;
    for (int k = 0; k < 3; ++k) w[k] = w[k] + c * q[i][k];
  }
}

/* mm10_h_voche (mm10_b.f:1896) */
static double h_voche(const cprops* p, const cstate* np1, const cstate* n,
                      const double* s, double tt) {
  double h = zero;
  for (int i = 0; i < NSL; ++i) {
    double gi = slipinc(p, np1, s, tt, i);
    double ht = one - (tt - p->tau_y) / p->tau_v + np1->tau_l[i] / (tt - p->tau_y);
    h = h + pow(fabs(ht), p->voce_m) * sgn(ht) * fabs(gi);
  }
  return n->tau_tilde + p->theta_0 * h;
}

/* mm10_formR1 (mm10_b.f:1073) */
static void formR1(const cprops* p, const cstate* np1, const cstate* n,
                   const double* s, double tt, double* R1) {
  double dbarp[6], wp[3], symTW[6], w1[6], w2[6];
  CNT(2);
  ++t_evals;
  form_dbarp(p, np1, s, tt, dbarp);
  form_w(p, np1, s, tt, np1->qc, wp);
  symsw(s, wp, symTW);
  for (int k = 0; k < 6; ++k) w1[k] = np1->D[k] - dbarp[k];
  mv6(w2, p->cr->C, w1);
  for (int k = 0; k < 6; ++k)
    R1[k] = s[k] - n->stress[k] - w2[k] + two * symTW[k];
}

/* mm10_formR (mm10_b.f:1037) with formR2 (:63), which also records the
   hardening rate */
static void formR(const cprops* p, cstate* np1, const cstate* n,
                  const double* x, double* R) {
  CNT(3);
  formR1(p, np1, n, x, x[6], R);
  double h = h_voche(p, np1, n, x, x[6]);
  R[6] = x[6] - h;
  np1->tt_rate = (h - n->tau_tilde) / np1->tinc;
}

/* mm10_formJ11 (mm10_b.f:177) for Voce */
static void formJ11(const cprops* p, const cstate* np1, const double* s,
                    double tt, double* J11) {
  double symtq[NSL][6], dgdt[NSL];
  CNT(4);
  memset(J11, 0, sizeof(double) * 36);
  for (int i = 0; i < NSL; ++i) symsw(s, np1->qc[i], symtq[i]);
  for (int i = 0; i < NSL; ++i) { /* mm10_dgdt_voche */
    double rs = rs_of(np1, s, i);
    double d = pow(fabs(rs), p->rate_n - one);
    d = np1->dg * p->rate_n / pow(tt, p->rate_n) * d;
    dgdt[i] = d + np1->tinc * p->iD_v;
  }
  for (int i = 0; i < NSL; ++i) {
    double wv[6];
    mv6_plus(wv, p->cr->C, np1->ms[i], symtq[i], two);
    for (int c = 0; c < 6; ++c) { /* DGER */
      double t = dgdt[i] * np1->ms[i][c];
      if (np1->ms[i][c] != 0.0)
        for (int r = 0; r < 6; ++r) J11[r + 6 * c] = J11[r + 6 * c] + wv[r] * t;
    }
  }
  double wp[3];
  form_w(p, np1, s, tt, np1->qc, wp);
  double iw[36];
  memset(iw, 0, sizeof iw); /* mm10_iw (mm10_b.f:1613) */
  M6(iw, 0, 3) = two * wp[2];
  M6(iw, 0, 5) = -two * wp[1];
  M6(iw, 1, 3) = two * wp[2];
  M6(iw, 1, 4) = -two * wp[0];
  M6(iw, 2, 4) = two * wp[0];
  M6(iw, 2, 5) = two * wp[1];
  M6(iw, 3, 0) = wp[2];
  M6(iw, 3, 1) = -wp[2];
  M6(iw, 3, 4) = -wp[1];
  M6(iw, 3, 5) = wp[0];
  M6(iw, 4, 1) = wp[0];
  M6(iw, 4, 2) = -wp[0];
  M6(iw, 4, 3) = wp[1];
  M6(iw, 4, 5) = wp[2];
  M6(iw, 5, 0) = wp[1];
  M6(iw, 5, 2) = -wp[1];
  M6(iw, 5, 3) = wp[0];
  M6(iw, 5, 4) = -wp[2];
  for (int k = 0; k < 36; ++k) J11[k] = J11[k] + iw[k];
  for (int i = 0; i < 6; ++i) J11[i + 6 * i] = J11[i + 6 * i] + one;
}

/* mm10_formJ (mm10_b.f:909): the 7x7 Jacobian (column-major, ld 7) */
static void formJ(const cprops* p, const cstate* np1, const double* x,
                  double* J) {
  const double* s = x;
  double tt = x[6];
  double J11[36];
  CNT(5);
  formJ11(p, np1, s, tt, J11);
  /* J12 (mm10_b.f:281) */
  double J12[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < NSL; ++i) {
    double st[6], tm[6];
    symsw(s, np1->qc[i], st);
    double dgam = slipinc(p, np1, s, tt, i);
    double dgdh = -p->rate_n / tt * dgam; /* mm10_dgdh_voche */
    mv6_plus(tm, p->cr->C, np1->ms[i], st, two);
    if (dgdh != 0.0)
      for (int r = 0; r < 6; ++r) J12[r] = J12[r] + tm[r] * (one * dgdh);
  }
  /* J21 = -estress (mm10_estress_voche, mm10_b.f:1925) */
  double et[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < NSL; ++i) {
    double rs = rs_of(np1, s, i);
    double ht = one - (tt - p->tau_y) / p->tau_v + np1->tau_l[i] / (tt - p->tau_y);
    double c = pow(fabs(ht), p->voce_m) * sgn(ht) *
               pow(fabs(rs), p->rate_n - two) * rs;
    for (int k = 0; k < 6; ++k) et[k] = et[k] + c * np1->ms[i][k];
  }
  double sc = p->theta_0 * np1->dg * p->rate_n / pow(tt, p->rate_n);
  for (int k = 0; k < 6; ++k) et[k] = sc * et[k];
  /* J22 = ehard (mm10_ehard_voche, mm10_b.f:1958) */
  double etau = zero;
  for (int i = 0; i < NSL; ++i) {
    double gi = slipinc(p, np1, s, tt, i);
    double ht = one - (tt - p->tau_y) / p->tau_v + np1->tau_l[i] / (tt - p->tau_y);
    double dty = tt - p->tau_y;
    etau = etau + (p->voce_m * (-one / p->tau_v - np1->tau_l[i] / (dty * dty)) *
                       fabs(gi) / fabs(ht) -
                   fabs(gi) * p->rate_n / tt * sgn(ht) * sgn(gi)) *
                      pow(fabs(ht), p->voce_m);
  }
  etau = p->theta_0 * etau;
  etau = one - etau;
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r) J[r + 7 * c] = J11[r + 6 * c];
  for (int r = 0; r < 6; ++r) J[r + 7 * 6] = J12[r];
  for (int c = 0; c < 6; ++c) J[6 + 7 * c] = -et[c];
  J[6 + 7 * 6] = etau;
}

/* ------------------------------------------------------------------ */
/* mm10_solve (mm10_a.f:2901): predictor then corrector                 */
/* ------------------------------------------------------------------ */
static int any_nan(const double* x, int n) {
  for (int i = 0; i < n; ++i)
    if (isnan(x[i])) return 1;
  return 0;
}

static void solve(const cprops* p, cstate* np1, const cstate* n,
                  double* stress, double* tt, int* fail, double dtinc,
                  double* J) {
  CNT(1);
  const double c = 1.0e-4, red = 0.5;
  const int mls = 10, mmin = 1;
  double x[7], inR1;
  for (int k = 0; k < 6; ++k) x[k] = stress[k];
  x[6] = *tt;
  /* ---- mm10_solve_predict ---- */
  {
    int iter = 0;
    double x1[6], x2, R1[6], d1[6], d2[6], t1, t2, dtrace;
    for (int k = 0; k < 6; ++k) x1[k] = x[k];
    double dt = dtinc;
    for (int k = 0; k < 6; ++k) d1[k] = np1->D[k];
    dtrace = (d1[0] + d1[1] + d1[2]) / three;
    d1[0] = d1[0] - dtrace; d1[1] = d1[1] - dtrace; d1[2] = d1[2] - dtrace;
    t1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    t2 = d1[3] * d1[3] + d1[4] * d1[4] + d1[5] * d1[5];
    if (t1 + t2 == zero)
      for (int k = 0; k < 6; ++k) d1[k] = zero;
    else
      for (int k = 0; k < 6; ++k) d1[k] = d1[k] / sqrt(t1 + t2);
    for (int k = 0; k < 6; ++k) d2[k] = n->D[k];
    dtrace = (d2[0] + d2[1] + d2[2]) / three;
    d2[0] = d2[0] - dtrace; d2[1] = d2[1] - dtrace; d2[2] = d2[2] - dtrace;
    t1 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    t2 = d2[3] * d2[3] + d2[4] * d2[4] + d2[5] * d2[5];
    if (t1 + t2 == zero)
      for (int k = 0; k < 6; ++k) d2[k] = zero;
    else
      for (int k = 0; k < 6; ++k) d2[k] = d2[k] / sqrt(t1 + t2);
    t1 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    t2 = d1[3] * d2[3] + d1[4] * d2[4] + d1[5] * d2[5];
    double cos_ang = t1 + t2 > zero ? t1 + t2 : zero;
    x2 = x[6] + cos_ang * n->tt_rate * dt;
    formR1(p, np1, n, x1, x2, R1);
    double nR1 = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] +
                      R1[3] * R1[3] + R1[4] * R1[4] + R1[5] * R1[5]);
    inR1 = nR1;
    int ok = 1;
    while (nR1 > p->atol1 && nR1 / inR1 > p->rtol1) {
      double J11[36], mJ[36], dx1[6], wv[6];
      formJ11(p, np1, x1, x2, J11);
      for (int k = 0; k < 6; ++k) dx1[k] = R1[k];
      for (int k = 0; k < 36; ++k) mJ[k] = -J11[k];
      lu_solve(6, mJ, 1, dx1);
      double alpha = one;
      double dotR1 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] +
                     R1[3] * R1[3] + R1[4] * R1[4] + R1[5] * R1[5];
      double ls1 = half * dotR1;
      for (int j = 0; j < 6; ++j) /* mm10_a_mult_type_2t */
        wv[j] = M6(J11, 0, j) * R1[0] + M6(J11, 1, j) * R1[1] +
                M6(J11, 2, j) * R1[2] + M6(J11, 3, j) * R1[3] +
                M6(J11, 4, j) * R1[4] + M6(J11, 5, j) * R1[5];
      double ls2 = c * (dx1[0] * wv[0] + dx1[1] * wv[1] + dx1[2] * wv[2] +
                        dx1[3] * wv[3] + dx1[4] * wv[4] + dx1[5] * wv[5]);
      int ls = 0;
      for (;;) {
        double nlsx = ls1 + ls2 * alpha, xn[6];
        for (int k = 0; k < 6; ++k) xn[k] = x1[k] + alpha * dx1[k];
        formR1(p, np1, n, xn, x2, R1);
        dotR1 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] +
                R1[3] * R1[3] + R1[4] * R1[4] + R1[5] * R1[5];
        nR1 = sqrt(dotR1);
        double nRs = half * dotR1;
        if (nRs <= nlsx || ls > mls) {
          for (int k = 0; k < 6; ++k) x1[k] = xn[k];
          break;
        }
        alpha = red * alpha;
        ls = ls + 1;
      }
      iter = iter + 1;
      if (iter > p->miter || any_nan(x1, 6)) {
        *fail = 1;
        ok = 0;
        break;
      }
    }
    if (ok) {
      for (int k = 0; k < 6; ++k) x[k] = x1[k];
      x[6] = x2;
    }
  }
  /* ---- mm10_solve_update ---- */
  {
    int iter = 0;
    double R[7];
    formR(p, np1, n, x, R);
    double dotR = 0.0;
    for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
    double nR = sqrt(dotR), inR = nR;
    if (inR == zero) inR = inR1;
    while ((nR > p->atol && nR / inR > p->rtol) || iter < mmin) {
      double mJ[49], dx[7], wv[7];
      formJ(p, np1, x, J);
      for (int k = 0; k < 7; ++k) dx[k] = R[k];
      for (int k = 0; k < 49; ++k) mJ[k] = -J[k];
      lu_solve(7, mJ, 1, dx);
      double alpha = one;
      dotR = 0.0;
      for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
      double ls1 = half * dotR;
      for (int j = 0; j < 7; ++j) { /* DGEMM('T','N') */
        double t = 0.0;
        for (int l = 0; l < 7; ++l) t = t + J[l + 7 * j] * R[l];
        wv[j] = t;
      }
      double dd = 0.0;
      for (int k = 0; k < 7; ++k) dd = dd + dx[k] * wv[k];
      double ls2 = c * dd;
      int ls = 0;
      for (;;) {
        double nlsx = ls1 + ls2 * alpha, xn[7];
        for (int k = 0; k < 7; ++k) xn[k] = x[k] + alpha * dx[k];
        formR(p, np1, n, xn, R);
        dotR = 0.0;
        for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
        nR = sqrt(dotR);
        double nRs = half * dotR;
        if (nRs <= nlsx || ls > mls) {
          for (int k = 0; k < 7; ++k) x[k] = xn[k];
          break;
        }
        alpha = red * alpha;
        ls = ls + 1;
      }
      iter = iter + 1;
      if (iter > p->miter || any_nan(x, 7)) {
        *fail = 1;
        break;
      }
    }
  }
  for (int k = 0; k < 6; ++k) stress[k] = x[k];
  *tt = x[6];
}

/* mm10_solve_strup (mm10_a.f:2657).  Returns the no-load flag. */
static int solve_strup(const cprops* p, cstate* np1, const cstate* n,
                       int* fail, int iter0_off, double* J) {
  const double mult = 0.5;
  const int mcuts = 4;
  double frac = zero, step = one;
  int cuts = 0;
  double stress[6], ostress[6], tt, ott;
  cstate curr;
  memset(&curr, 0, sizeof curr);
  for (int k = 0; k < 6; ++k) stress[k] = ostress[k] = n->stress[k];
  tt = ott = n->tau_tilde;
  setup(p, np1, n);
  *fail = 0;
  double temp1 = np1->D[0] * np1->D[0] + np1->D[1] * np1->D[1] +
                 np1->D[2] * np1->D[2] + np1->D[3] * np1->D[3] +
                 np1->D[4] * np1->D[4] + np1->D[5] * np1->D[5];
  double temp2 = stress[0] * stress[0] + stress[1] * stress[1] +
                 stress[2] * stress[2] + stress[3] * stress[3] +
                 stress[4] * stress[4] + stress[5] * stress[5];
  int no_load = temp2 == zero && temp1 == zero;
  if (iter0_off || no_load) { /* mm10_solve_strup_a */
    memcpy(np1->tangent, p->cr->C, sizeof np1->tangent);
    if (!no_load) {
      double R1[6];
      formR1(p, np1, n, stress, tt, R1);
      for (int k = 0; k < 6; ++k) stress[k] = stress[k] - R1[k];
    }
    curr.tt_rate = zero;
  } else { /* mm10_solve_strup_iterate */
    while (frac < one) {
      double Dw[6];
      for (int k = 0; k < 6; ++k) Dw[k] = np1->D[k] * (step + frac);
      double tinc_w = np1->tinc * (step + frac);
      setup_np1(np1->R, Dw, tinc_w, &curr);
      setup(p, &curr, n);
      tt = n->tau_tilde;
      solve(p, &curr, n, stress, &tt, fail, np1->tinc * step, J);
      if (*fail) {
        for (int k = 0; k < 6; ++k) stress[k] = ostress[k];
        tt = ott;
        step = step * mult;
        cuts = cuts + 1;
        if (cuts > mcuts) break;
        *fail = 0;
      } else {
        for (int k = 0; k < 6; ++k) ostress[k] = stress[k];
        ott = tt;
        frac = frac + step;
      }
    }
    int failed = *fail || isnan(tt) || any_nan(stress, 6);
    if (failed) {
      *fail = 1;
      for (int k = 0; k < 6; ++k) np1->stress[k] = n->stress[k];
      np1->tau_tilde = n->tau_tilde;
      return no_load;
    }
  }
  for (int k = 0; k < 6; ++k) np1->stress[k] = stress[k];
  np1->tau_tilde = tt;
  np1->tt_rate = curr.tt_rate;
  return no_load;
}

/* mm10_tangent (mm10_a.f:664) for Voce: ed = 0 and d(gamma)/dD = 0, so
   JA = JB = 0 and T = JJ^-1 C with JJ = J11 - J12 J22^-1 J21 */
static void tangent(const cprops* p, cstate* np1, const double* J) {
  double JJ[36], beta[12];
  for (int c = 0; c < 6; ++c)
    for (int r = 0; r < 6; ++r) JJ[r + 6 * c] = J[r + 7 * c];
  double alpha = J[6 + 7 * 6];
  for (int c = 0; c < 6; ++c) {
    beta[c] = J[6 + 7 * c];
    beta[6 + c] = zero;
  }
  double a1 = alpha;
  lu_solve(1, &a1, 12, beta);
  for (int j = 0; j < 6; ++j) { /* DGEMM: JJ = JJ - J12 beta(:,1:6) */
    double t = -one * beta[j];
    if (t != 0.0)
      for (int i = 0; i < 6; ++i) JJ[i + 6 * j] = JJ[i + 6 * j] + t * J[i + 7 * 6];
  }
  double JR[36];
  for (int k = 0; k < 36; ++k) JR[k] = p->cr->C[k] - zero - zero;
  lu_solve(6, JJ, 6, JR);
  memcpy(np1->tangent, JR, sizeof JR);
}

/* mm10_a_make_symm_1 */
static void make_symm(double* a) {
  double w[36];
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) w[i + 6 * j] = (a[i + 6 * j] + a[j + 6 * i]) * half;
  memcpy(a, w, sizeof w);
}

/* mm10_update_rotation (mm10_a.f:3375) */
static void update_rotation(const cprops* p, cstate* np1, const cstate* n) {
  double wbarp[3], expw[9];
  form_w(p, np1, np1->stress, np1->tau_tilde, np1->qs, wbarp);
  orc_mm10_expmw(wbarp, expw);
  mul33(np1->Rp, expw, n->Rp);
}

static double atan2_02pi(double a, double b) {
  double v = atan2(a, b);
  if (v < zero) v = v + two * pi;
  return v;
}

/* mm10_output (mm10_a.f:3493) with mm10_update_euler_angles (:1195) */
static void output(const cprops* p, cstate* np1, double* p_strain_ten_c) {
  const int nslip = NSL;
  for (int i = 0; i < nslip; ++i)
    np1->slip_incs[i] = slipinc(p, np1, np1->stress, np1->tau_tilde, i);
  { /* Euler angles, Kocks convention */
    double work1[9], full[9];
    for (int i = 0; i < 3; ++i) /* mm10_a_mult_type_3t: Rp R^T */
      for (int j = 0; j < 3; ++j)
        M3(work1, i, j) = M3(np1->Rp, i, 0) * M3(np1->R, j, 0) +
                          M3(np1->Rp, i, 1) * M3(np1->R, j, 1) +
                          M3(np1->Rp, i, 2) * M3(np1->R, j, 2);
    mul33(full, p->cr->g, work1);
    double psiK = atan2_02pi(M3(full, 2, 1), M3(full, 2, 0));
    double phiK = atan2_02pi(M3(full, 1, 2), M3(full, 0, 2));
    if (M3(full, 2, 2) > one) M3(full, 2, 2) = one;
    double thetaK = acos(M3(full, 2, 2));
    if (p->angle_type == 1) {
      np1->euler[0] = one_eighty / pi * psiK;
      np1->euler[1] = one_eighty / pi * thetaK;
      np1->euler[2] = one_eighty / pi * phiK;
    } else {
      np1->euler[0] = psiK;
      np1->euler[1] = thetaK;
      np1->euler[2] = phiK;
    }
  }
  double dif_slp[NSL], ed[6], t1, t2;
  for (int i = 0; i < nslip; ++i) {
    double rs = rs_of(np1, np1->stress, i);
    dif_slp[i] = rs * np1->tinc * p->iD_v;
  }
  for (int k = 0; k < 6; ++k) ed[k] = zero;
  for (int i = 0; i < nslip; ++i)
    for (int k = 0; k < 6; ++k) ed[k] = ed[k] + dif_slp[i] * np1->ms[i][k];
  for (int k = 0; k < 6; ++k) np1->ed[k] = ed[k] / np1->tinc;
  t1 = ed[0] * ed[0] + ed[1] * ed[1] + ed[2] * ed[2];
  t2 = ed[3] * ed[3] + ed[4] * ed[4] + ed[5] * ed[5];
  double ed_dot = sqrt(two / three * (t1 + half * t2));
  ed_dot = ed_dot / np1->tinc;
  np1->u[14] = ed_dot;
  np1->work_inc = np1->stress[0] * np1->D[0] + np1->stress[1] * np1->D[1] +
                  np1->stress[2] * np1->D[2] + np1->stress[3] * np1->D[3] +
                  np1->stress[4] * np1->D[4] + np1->stress[5] * np1->D[5];
  double eeunrot[6], erot[36], ee[6], dbarp[6], wp[3], ewwe[6], ep[6];
  for (int k = 0; k < 6; ++k) eeunrot[k] = np1->stress[k];
  chol_solve6(p->cr->C, eeunrot);
  orc_mm10_rt2rve(np1->R, erot);
  mv6(ee, erot, eeunrot);
  for (int k = 0; k < 6; ++k) dbarp[k] = zero;
  for (int i = 0; i < nslip; ++i)
    for (int k = 0; k < 6; ++k)
      dbarp[k] = dbarp[k] + np1->slip_incs[i] * np1->ms[i][k];
  form_w(p, np1, np1->stress, np1->tau_tilde, np1->qc, wp);
  symsw(ee, wp, ewwe);
  for (int k = 0; k < 6; ++k) ep[k] = dbarp[k] + ewwe[k];
  for (int k = 0; k < 6; ++k) np1->ep[k] = ep[k] / np1->tinc;
  t1 = ep[0] * ep[0] + ep[1] * ep[1] + ep[2] * ep[2];
  t2 = ep[3] * ep[3] + ep[4] * ep[4] + ep[5] * ep[5];
  double ep_dot = sqrt(two / three * (t1 + half * t2));
  ep_dot = ep_dot / np1->tinc;
  np1->u[10] = ep_dot;
  for (int k = 0; k < 6; ++k) ep[k] = ep[k] + ed[k];
  for (int k = 0; k < 6; ++k) p_strain_ten_c[k] = ep[k];
  t1 = ep[0] * ep[0] + ep[1] * ep[1] + ep[2] * ep[2];
  t2 = ep[3] * ep[3] + ep[4] * ep[4] + ep[5] * ep[5];
  np1->p_strain_inc = sqrt(two / three * (t1 + half * t2));
  np1->p_work_inc = np1->stress[0] * ep[0] + np1->stress[1] * ep[1] +
                    np1->stress[2] * ep[2] + np1->stress[3] * ep[3] +
                    np1->stress[4] * ep[4] + np1->stress[5] * ep[5];
  for (int k = 0; k < 6; ++k) np1->eps[k] = ee[k];
  double ec_dot = np1->p_strain_inc / np1->tinc, n_eff;
  if (ec_dot > zero) {
    for (int k = 0; k < 6; ++k) ep[k] = ep[k] / np1->tinc;
    double dgdt[NSL];
    for (int i = 0; i < nslip; ++i) {
      double rs = rs_of(np1, np1->stress, i);
      double d = pow(fabs(rs), p->rate_n - one);
      d = np1->dg * p->rate_n / pow(np1->tau_tilde, p->rate_n) * d;
      dgdt[i] = d + np1->tinc * p->iD_v;
    }
    n_eff = zero;
    for (int i = 0; i < nslip; ++i) {
      double rs = rs_of(np1, np1->stress, i);
      const double* m = np1->ms[i];
      t1 = m[0] * ep[0] + m[1] * ep[1] + m[2] * ep[2];
      t2 = m[3] * ep[3] + m[4] * ep[4] + m[5] * ep[5];
      double ec_slip = t1 + half * t2;
      t1 = rs * dgdt[i] / ec_dot;
      t2 = ec_slip / ec_dot / np1->tinc;
      n_eff = n_eff + (two / three) * t1 * t2;
    }
  } else {
    n_eff = 1.0e10;
  }
  np1->u[11] = n_eff;
  double s_trace = (np1->stress[0] + np1->stress[1] + np1->stress[2]) / three;
  double se[6];
  for (int k = 0; k < 6; ++k) se[k] = np1->stress[k];
  se[0] = se[0] - s_trace;
  se[1] = se[1] - s_trace;
  se[2] = se[2] - s_trace;
  t1 = se[0] * se[0] + se[1] * se[1] + se[2] * se[2];
  t2 = se[3] * se[3] + se[4] * se[4] + se[5] * se[5];
  s_trace = sqrt(onept5 * (t1 + two * t2));
  np1->u[12] = s_trace;
  double B_eff;
  if (ec_dot < 1.0e-100)
    B_eff = zero;
  else if (n_eff > 100.0)
    B_eff = -one;
  else
    B_eff = ec_dot / pow(s_trace, n_eff);
  np1->u[13] = B_eff;
  for (int i = 0; i < nslip; ++i)
    np1->slip_incs[i] = np1->slip_incs[i] + dif_slp[i];
  double maxslip = zero;
  int sysID = 0;
  for (int i = 0; i < nslip; ++i) {
    double cs = fabs(np1->slip_incs[i]);
    if (cs > maxslip) {
      maxslip = cs;
      sysID = i + 1;
    }
  }
  np1->u[5] = maxslip / np1->tinc;
  np1->u[6] = (double)sysID;
  int numAct = 0;
  maxslip = pt1 * maxslip;
  for (int i = 0; i < nslip; ++i)
    if (fabs(np1->slip_incs[i]) >= maxslip) numAct = numAct + 1;
  np1->u[7] = (double)numAct;
}

/* ------------------------------------------------------------------ */
/* the Gauss-point driver: mm10 + mm10_a_do_crystal + store             */
/* ------------------------------------------------------------------ */
int orc_mm10_hist_size(int ncry) { return NCOMMON + NCRYS * ncry; }

int orc_mm10_gp(const w3d_material* mat, const orc_crystal* crys, int step,
                int iter0_off, double dt, const double* uddt,
                const double* urcs_n, double* urcs_n1, double* hn,
                double* hn1, int asym, const double* rn1) {
  const int ncry = mat->ncrystals > 0 ? mat->ncrystals : 1;
  static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  const double* Rur = rn1 ? rn1 : I3;
  if (step == 1) { /* mm10_init_general_hist, _uout_hist, _slip_hist */
    for (int k = 0; k < 63; ++k) hn[k] = zero;
    for (int k = 0; k < 9; ++k) hn[63 + k] = I3[k];
    for (int k = 72; k < NCOMMON; ++k) hn[k] = zero;
  }
  double sig_avg[6] = {0, 0, 0, 0, 0, 0}, tang_avg[36], slip_avg[NSL];
  double t_work_inc = zero, p_work_inc = zero, p_strain_inc = zero;
  memset(tang_avg, 0, sizeof tang_avg);
  memset(slip_avg, 0, sizeof slip_avg);
  for (int c = 0; c < ncry; ++c) {
    cprops p;
    p.cr = &crys[c];
    p.rate_n = mat->harden_n;
    p.tau_y = mat->tau_y;
    p.tau_v = mat->tau_v;
    p.voce_m = mat->voce_m;
    p.theta_0 = mat->theta_0;
    p.iD_v = mat->iD_v;
    p.eps_dot_0_y = mat->eps_dot_0_y;
    p.atol = mat->atol;
    p.atol1 = mat->atol1;
    p.rtol = mat->rtol;
    p.rtol1 = mat->rtol1;
    p.miter = mat->miter;
    p.alter_mode = mat->alter_mode;
    p.angle_type = mat->angle_type;
    double* h = hn + NCOMMON + NCRYS * c;
    if (step == 1) { /* mm10_init_cc_hist0 + mm10_init_voche */
      for (int k = 0; k < 6; ++k) h[k] = urcs_n[k];
      for (int k = 0; k < 3; ++k) h[6 + k] = crys[c].angles[k];
      for (int k = 0; k < 9; ++k) h[9 + k] = I3[k];
      for (int k = 18; k < 42; ++k) h[k] = zero;
      h[42] = p.tau_y + 1.0e-5;
      for (int k = 43; k < 58; ++k) h[k] = zero;
    }
    /* mm10_copy_cc_hist */
    cstate n, np1;
    memset(&n, 0, sizeof n);
    memset(&np1, 0, sizeof np1);
    for (int k = 0; k < 9; ++k) n.R[k] = hn[63 + k];
    for (int k = 0; k < 6; ++k) n.stress[k] = h[k];
    for (int k = 0; k < 3; ++k) n.euler[k] = h[6 + k];
    for (int k = 0; k < 9; ++k) n.Rp[k] = h[9 + k];
    for (int k = 0; k < 6; ++k) n.D[k] = h[18 + k];
    for (int k = 0; k < 6; ++k) n.eps[k] = h[24 + k];
    for (int k = 0; k < NSL - 1; ++k) n.slip_incs[k] = h[30 + k];
    n.tau_tilde = h[42];
    for (int k = 0; k < 14; ++k) n.u[k] = h[43 + k];
    n.tt_rate = h[58];
    for (int k = 0; k < 6; ++k) n.ep[k] = h[59 + k];
    for (int k = 0; k < 6; ++k) n.ed[k] = h[65 + k];
    /* mm10_setup_np1 with R n+1 (identity for small strains) */
    setup_np1(Rur, uddt, dt, &np1);
    /* mm10_solve_crystal */
    double J[49], pst[6];
    int fail = 0;
    memset(J, 0, sizeof J);
    t_evals = 0;
    int no_load = solve_strup(&p, &np1, &n, &fail, iter0_off, J);
    if (!(iter0_off || no_load)) {
      CNT(0);
      __atomic_fetch_add(&g_hist[t_evals < 63 ? t_evals : 63], 1,
                         __ATOMIC_RELAXED);
    }
    if (fail) return 1;
    if (!(iter0_off || no_load)) {
      tangent(&p, &np1, J);
      /* mm10_a.f:1165: kept as formed for asymmetric assembly */
      if (!asym) make_symm(np1.tangent);
      update_rotation(&p, &np1, &n);
      output(&p, &np1, pst);
    }
    for (int k = 0; k < 6; ++k) sig_avg[k] = sig_avg[k] + np1.stress[k];
    for (int k = 0; k < 36; ++k) tang_avg[k] = tang_avg[k] + np1.tangent[k];
    for (int k = 0; k < NSL; ++k) slip_avg[k] = slip_avg[k] + np1.slip_incs[k];
    t_work_inc = t_work_inc + np1.work_inc;
    p_work_inc = p_work_inc + np1.p_work_inc;
    p_strain_inc = p_strain_inc + np1.p_strain_inc;
    /* mm10_store_cryhist */
    double* h1 = hn1 + NCOMMON + NCRYS * c;
    for (int k = 0; k < 6; ++k) h1[k] = np1.stress[k];
    for (int k = 0; k < 3; ++k) h1[6 + k] = np1.euler[k];
    for (int k = 0; k < 9; ++k) h1[9 + k] = np1.Rp[k];
    for (int k = 0; k < 6; ++k) h1[18 + k] = np1.D[k];
    for (int k = 0; k < 6; ++k) h1[24 + k] = np1.eps[k];
    for (int k = 0; k < NSL; ++k) h1[30 + k] = np1.slip_incs[k];
    h1[42] = np1.tau_tilde;
    for (int k = 0; k < 15; ++k) h1[43 + k] = np1.u[k];
    h1[58] = np1.tt_rate;
    for (int k = 0; k < 6; ++k) h1[59 + k] = np1.ep[k];
    for (int k = 0; k < 6; ++k) h1[65 + k] = np1.ed[k];
  }
  /* mm10_a_crystal_avgs */
  double rncry = (double)ncry;
  for (int k = 0; k < 6; ++k) sig_avg[k] = sig_avg[k] / rncry;
  for (int k = 0; k < 36; ++k) tang_avg[k] = tang_avg[k] / rncry;
  for (int k = 0; k < NSL; ++k) slip_avg[k] = slip_avg[k] / rncry;
  t_work_inc = t_work_inc / rncry;
  p_work_inc = p_work_inc / rncry;
  p_strain_inc = p_strain_inc / rncry;
  /* mm10_a_store_crystal */
  for (int k = 0; k < 6; ++k) urcs_n1[k] = sig_avg[k];
  for (int k = 0; k < 36; ++k) hn1[k] = tang_avg[k];
  for (int k = 36; k < 63; ++k) hn1[k] = zero;
  for (int k = 0; k < 9; ++k) hn1[63 + k] = Rur[k];
  for (int k = 0; k < NSL; ++k) hn1[75 + k] = hn[75 + k] + slip_avg[k];
  urcs_n1[6] = urcs_n[6] + t_work_inc;
  urcs_n1[7] = urcs_n[7] + p_work_inc;
  urcs_n1[8] = urcs_n[8] + p_strain_inc;
  hn1[72] = hn[72] + t_work_inc;
  hn1[73] = hn[73] + p_work_inc;
  hn1[74] = hn[74] + p_strain_inc;
  return 0;
}

/* Residual and Jacobian of the local problem at x = [stress; tau_tilde]
   for a crystal state built like mm10_solve_strup's first sub-step:
   exported for the finite-difference pins of formJ. */
void orc_mm10_residual(const w3d_material* mat, const orc_crystal* cr,
                       const double* hist_crystal, const double* uddt,
                       double dt, const double* x, double* R, double* J,
                       const double* rn1) {
  cprops p;
  p.cr = cr;
  p.rate_n = mat->harden_n;
  p.tau_y = mat->tau_y;
  p.tau_v = mat->tau_v;
  p.voce_m = mat->voce_m;
  p.theta_0 = mat->theta_0;
  p.iD_v = mat->iD_v;
  p.eps_dot_0_y = mat->eps_dot_0_y;
  p.alter_mode = mat->alter_mode;
  static const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  cstate n, np1;
  memset(&n, 0, sizeof n);
  memcpy(n.R, I3, sizeof I3);
  for (int k = 0; k < 6; ++k) n.stress[k] = hist_crystal[k];
  for (int k = 0; k < 9; ++k) n.Rp[k] = hist_crystal[9 + k];
  n.tau_tilde = hist_crystal[42];
  setup_np1(rn1 ? rn1 : I3, uddt, dt, &np1);
  setup(&p, &np1, &n);
  double xx[7];
  memcpy(xx, x, sizeof xx);
  if (R) formR(&p, &np1, &n, xx, R);
  if (J) formJ(&p, &np1, xx, J);
}

/* ------------------------------------------------------------------ */
/* grad Fe of single-crystal elements: rknstr_finish_cp (rknstr.f)     */
/* ------------------------------------------------------------------ */
/* Least-squares solution of min |A x - b| for A (m x 4, column-major,
   leading dimension m), as LAPACK DGELS('N') takes it: Householder QR
   (dgeqr2/dlarfg) and back substitution when m >= 4; for m < 4 the
   minimum-norm solution A^T (A A^T)^-1 b (m = 1 here: the one-point
   rule).  x receives 4 values. */
void orc_lsq4(int m, const double* a_in, const double* b_in, double* x) {
  double a[4 * MXGP], b[MXGP];
  memcpy(a, a_in, sizeof(double) * 4 * m);
  memcpy(b, b_in, sizeof(double) * m);
  if (m < 4) { /* only the one-point rule reaches here */
    double s = 0.0;
    for (int j = 0; j < 4; ++j) s = s + a[j * m] * a[j * m];
    for (int j = 0; j < 4; ++j) x[j] = s > 0.0 ? a[j * m] * b[0] / s : 0.0;
    return;
  }
  for (int j = 0; j < 4; ++j) {
    double* col = a + j * m;
    double alpha = col[j], xn = 0.0;
    for (int i = j + 1; i < m; ++i) xn = xn + col[i] * col[i];
    xn = sqrt(xn);
    double tau = 0.0;
    if (xn != 0.0) {
      double beta = -copysign(sqrt(alpha * alpha + xn * xn), alpha);
      tau = (beta - alpha) / beta;
      double sc = 1.0 / (alpha - beta);
      for (int i = j + 1; i < m; ++i) col[i] = col[i] * sc;
      col[j] = beta;
    }
    /* apply H = I - tau v v^T (v = [1; col(j+1:m)]) to the remaining
       columns and to b */
    for (int k = j + 1; k <= 4; ++k) {
      double* y = k < 4 ? a + k * m : b;
      double w = y[j];
      for (int i = j + 1; i < m; ++i) w = w + col[i] * y[i];
      w = tau * w;
      y[j] = y[j] - w;
      for (int i = j + 1; i < m; ++i) y[i] = y[i] - w * col[i];
    }
  }
  for (int j = 3; j >= 0; --j) {
    double s = b[j];
    for (int k = j + 1; k < 4; ++k) s = s - a[k * m + j] * x[k];
    x[j] = s / a[j * m + j];
  }
}

/* mm10_calc_grads (mm10_a.f:1579) as rknstr_finish_cp calls it for small
   strain: Rt3d(3,3,ngp) is a flat copy (mm10_a_copy_vector) of 9 ngp words
   of the block's gathered n-history elem_hist(span, hist, ngp), starting
   at (element 1, first Rp word of crystal 1, point 1) -- words of several
   elements of the block, not element i's Rp.  For each (a, b) the
   coefficients of a linear fit Rt3d(a,b,p) ~ c1 xi + c2 eta + c3 zeta + c4
   over the Gauss points (DGELS) give grads(a,b,1:3).  hloc is the block's
   n-history in the (hist, ngp, span) layout; grads_vec receives grads in
   Fortran order (a fastest). */
int orc_mm10_calc_grads_geonl(int etype, int order, int ngp, int hist,
                              const double* hel, const double* rot,
                              const double jac[9], double* grads_vec) {
  /* mm10_a_invert_33 of the element's Jacobian */
  const double* a = jac;
#define A3(i, j) a[3 * ((i) - 1) + (j) - 1]
  double det = A3(1, 1) * A3(2, 2) * A3(3, 3) - A3(1, 1) * A3(2, 3) * A3(3, 2) -
               A3(1, 2) * A3(2, 1) * A3(3, 3) + A3(1, 2) * A3(2, 3) * A3(3, 1) +
               A3(1, 3) * A3(2, 1) * A3(3, 2) - A3(1, 3) * A3(2, 2) * A3(3, 1);
  double ji[9]; /* jacinv, row-major */
  int bad = fabs(det) <= 1.0e-10;
  if (bad) {
    for (int k = 0; k < 9; ++k) ji[k] = zero;
  } else {
    double cf[3][3];
    cf[0][0] = +(A3(2, 2) * A3(3, 3) - A3(2, 3) * A3(3, 2));
    cf[0][1] = -(A3(2, 1) * A3(3, 3) - A3(2, 3) * A3(3, 1));
    cf[0][2] = +(A3(2, 1) * A3(3, 2) - A3(2, 2) * A3(3, 1));
    cf[1][0] = -(A3(1, 2) * A3(3, 3) - A3(1, 3) * A3(3, 2));
    cf[1][1] = +(A3(1, 1) * A3(3, 3) - A3(1, 3) * A3(3, 1));
    cf[1][2] = -(A3(1, 1) * A3(3, 2) - A3(1, 2) * A3(3, 1));
    cf[2][0] = +(A3(1, 2) * A3(2, 3) - A3(1, 3) * A3(2, 2));
    cf[2][1] = -(A3(1, 1) * A3(2, 3) - A3(1, 3) * A3(2, 1));
    cf[2][2] = +(A3(1, 1) * A3(2, 2) - A3(1, 2) * A3(2, 1));
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) ji[3 * i + j] = cf[j][i] / det;
  }
#undef A3
  /* Rt3d(:, :, p) = Rp R^T (mm10_a_mult_type_3t), column-major */
  double rt[9 * MXGP];
  for (int p = 0; p < ngp; ++p) {
    const double* b = hel + (int64_t)p * hist + NCOMMON + 9; /* Rp, crys 1 */
    const double* c = rot + 9 * p;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        rt[9 * p + i + 3 * j] = M3(b, i, 0) * M3(c, j, 0) +
                                M3(b, i, 1) * M3(c, j, 1) +
                                M3(b, i, 2) * M3(c, j, 2);
  }
  double amat[4 * MXGP];
  for (int p = 0; p < ngp; ++p) {
    double xi, eta, zeta, w;
    orc_getgpts(etype, order, p + 1, &xi, &eta, &zeta, &w);
    amat[p] = xi;
    amat[ngp + p] = eta;
    amat[2 * ngp + p] = zeta;
    amat[3 * ngp + p] = 1.0;
  }
  for (int ia = 0; ia < 3; ++ia)
    for (int ib = 0; ib < 3; ++ib) {
      double rhs[MXGP], x[4];
      for (int p = 0; p < ngp; ++p) rhs[p] = rt[9 * p + ia + 3 * ib];
      orc_lsq4(ngp, amat, rhs, x);
      /* take to the current coordinates: jacinv (slopes) */
      for (int d = 0; d < 3; ++d)
        grads_vec[ia + 3 * ib + 9 * d] =
            ji[3 * d] * x[0] + ji[3 * d + 1] * x[1] + ji[3 * d + 2] * x[2];
    }
  return bad;
}

void orc_mm10_calc_grads(int etype, int order, int ngp, int span, int hist,
                         const double* hloc, double* grads_vec) {
  double amat[4 * MXGP];
  for (int p = 0; p < ngp; ++p) {
    double xi, eta, zeta, w;
    orc_getgpts(etype, order, p + 1, &xi, &eta, &zeta, &w);
    amat[p] = xi;
    amat[ngp + p] = eta;
    amat[2 * ngp + p] = zeta;
    amat[3 * ngp + p] = 1.0;
  }
  const int64_t l0 = (int64_t)(NCOMMON + 9) * span;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      double rhs[MXGP], x[4];
      for (int p = 0; p < ngp; ++p) {
        int64_t l = l0 + a + 3 * b + 9 * p;
        int64_t ie = l % span, wd = (l / span) % hist, g = l / ((int64_t)span * hist);
        rhs[p] = hloc[(ie * ngp + g) * hist + wd];
      }
      orc_lsq4(ngp, amat, rhs, x);
      for (int d = 0; d < 3; ++d) grads_vec[a + 3 * b + 9 * d] = x[d];
    }
}
