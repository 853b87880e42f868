/*
 * orc_mm01.c -- oracle: mm01 bilinear Mises model and its consistent
 * tangent cnst1, for one Gauss point of one element.
 *
 * TEST INFRASTRUCTURE ONLY (see w3d_oracle.h).  Restates mm01.f:24 (mm01),
 * :235 (mm01_set_history), :308 (mm01_init), :776 (mm01_general), :600
 * (mm01_sig_final), :524 (mm01_wrapup) and :1196 (cnst1), with the
 * reference's own truncated constants.  e, nu are the point's moduli at n+1
 * and e_n, nu_n those at n (they differ for temperature-dependent curves;
 * the elastic strain at n uses the latter, mm01.f:374-383).
 *
 * History words (1-based as in mm01_set_history): 1 lambda*dt, 2 k (yield
 * radius), 3 accumulated plastic strain, 4 state (an integer kept in the low
 * 32 bits of the double word, high bits zero), 5 H' at n+1, 6-11 back stress.
 * Stress words: 1-6 sigma, 7 energy density, 8 plastic work density,
 * 9 accumulated plastic strain increment sum.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "w3d_oracle.h"

static double state_word(int32_t st) {
  union { double d; int32_t i[2]; } u;
  u.i[0] = st;
  u.i[1] = 0;
  return u.d;
}
static int32_t state_of(double d) {
  union { double d; int32_t i[2]; } u;
  u.d = d;
  return u.i[0];
}

float orc_hprime(float e, float et) {
  volatile float num = e * et; /* REAL arithmetic, elprp.f:872 */
  volatile float den = e - et;
  return num / den;
}

int orc_mm01_gp(int step, double e, double nu, double beta, double hprime,
                int lnelas, double yld, double* cgn, double* cgn1,
                const double* deps, double* history, double* history1,
                double first_hist11, double* cep) {
  return orc_mm01_gp_t(step, e, nu, e, nu, beta, hprime, lnelas, yld, cgn,
                       cgn1, deps, history, history1, first_hist11, cep);
}

int orc_mm01_gp_t(int step, double e, double nu, double e_nv, double nu_nv,
                  double beta, double hprime, int lnelas, double yld,
                  double* cgn, double* cgn1, const double* deps,
                  double* history, double* history1, double first_hist11,
                  double* cep) {
  const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, half = 0.5,
               six = 6.0, four = 4.0;
  double* h = history;   /* 11 words, history(i,1..11) */
  double* h1 = history1;
  /* ---- mm01_set_history (step 1 only) ---- */
  if (step == 1) {
    const double root3 = sqrt(3.0);
    double kn = yld / root3;
    h[0] = zero; h[1] = kn; h[2] = zero; h[3] = state_word(3);
    h[4] = hprime;
    for (int k = 5; k < 11; ++k) h[k] = zero;
    memcpy(h1, h, sizeof(double) * 11);
    cgn[7] = zero;
    cgn[8] = zero;
    first_hist11 = zero;
  }
  /* ---- mm01_init ---- */
  const double htol = 0.000001, root3 = sqrt(3.0), root2 = sqrt(2.0),
               yld_tol = 0.0000001;
  double deps_vol = deps[0] + deps[1] + deps[2];
  double eps_mean = deps_vol / three;
  double de1 = deps[0] - eps_mean, de2 = deps[1] - eps_mean,
         de3 = deps[2] - eps_mean, de4 = deps[3], de5 = deps[4],
         de6 = deps[5];
  double e_n = e_nv, nu_n = nu_nv;
  double g_n = e_n / two / (one + nu_n);
  double een1 = (cgn[0] - nu_n * (cgn[1] + cgn[2])) / e_n;
  double een2 = (cgn[1] - nu_n * (cgn[0] + cgn[2])) / e_n;
  double een3 = (cgn[2] - nu_n * (cgn[0] + cgn[1])) / e_n;
  double een4 = cgn[3] / g_n, een5 = cgn[4] / g_n, een6 = cgn[5] / g_n;
  double eps_vol_n1 = een1 + een2 + een3 + deps_vol;
  double eps_mean_n = (een1 + een2 + een3) / three;
  double e1 = (een1 - eps_mean_n) + de1;
  double e2 = (een2 - eps_mean_n) + de2;
  double e3 = (een3 - eps_mean_n) + de3;
  double e4 = een4 + de4, e5 = een5 + de5, e6 = een6 + de6;
  double shear_mod = e / (two * (one + nu));
  double dse[6];
  dse[0] = two * shear_mod * e1;
  dse[1] = two * shear_mod * e2;
  dse[2] = two * shear_mod * e3;
  dse[3] = shear_mod * e4;
  dse[4] = shear_mod * e5;
  dse[5] = shear_mod * e6;
  double alpha_n[6];
  for (int k = 0; k < 6; ++k) alpha_n[k] = h[5 + k];
  double hbari_np1 = beta * hprime;
  double hbark_np1 = (one - beta) * hprime;
  double hbark_n = (one - beta) * h[4];
  double kbar = (yld + hbari_np1 * h[2]) / root3;
  double lk = one;
  if (fabs(hbark_n) > htol) lk = hbark_np1 / hbark_n;
  double rtse[6];
  for (int k = 0; k < 6; ++k) rtse[k] = dse[k] - alpha_n[k] * lk;
  double mrts = sqrt(rtse[0] * rtse[0] + rtse[1] * rtse[1] +
                     rtse[2] * rtse[2] +
                     two * (rtse[3] * rtse[3] + rtse[4] * rtse[4] +
                            rtse[5] * rtse[5]));
  double yf = mrts - root2 * kbar;
  for (int k = 0; k < 6; ++k) rtse[k] = dse[k] - alpha_n[k];
  mrts = sqrt(rtse[0] * rtse[0] + rtse[1] * rtse[1] + rtse[2] * rtse[2] +
              two * (rtse[3] * rtse[3] + rtse[4] * rtse[4] +
                     rtse[5] * rtse[5]));
  int instat = 3, yield = 0;
  if (!lnelas && yf >= yld_tol * root2 * kbar) {
    yield = 1;
    instat = 1;
  }
  double devstr[6];
  /* ---- mm01_general ---- */
  if (yield) {
    const double r2 = 1.414213562373095, twthrd = 0.666666666666667,
                 root23 = 0.816496580927, root2o3 = 0.4714045207910;
    double hbari_n = beta * h[4];
    (void)hbari_n;
    double a = (two * shear_mod + twthrd * hbark_np1) / mrts;
    double b = one - lk;
    double d = root2o3 * hbari_np1;
    double vbar = mrts * mrts;
    double wbar = h[5] * h[5] + h[6] * h[6] + h[7] * h[7] +
                  two * (h[8] * h[8] + h[9] * h[9] +
                         first_hist11 * first_hist11);
    double vwbar = rtse[0] * h[5] + rtse[1] * h[6] + rtse[2] * h[7] +
                   two * (rtse[3] * h[8] + rtse[4] * h[9] + rtse[5] * h[10]);
    double t1 = a * a * vbar - two * d * d;
    double t2 = -four * d * kbar - two * a * vbar - two * a * b * vwbar;
    double t3 = b * b * wbar + two * b * vwbar + vbar - two * kbar * kbar;
    double discr = t2 * t2 - four * t1 * t3;
    double lambda_deltat;
    if (discr < zero) {
      /* unreachable for isothermal data (lk = 1): see the design notes */
      lambda_deltat = zero;
    } else {
      double qroot1 = (-t2 + sqrt(discr)) / (two * t1);
      double qroot2 = (-t2 - sqrt(discr)) / (two * t1);
      lambda_deltat = qroot1 < qroot2 ? qroot1 : qroot2;
    }
    if (lambda_deltat < zero) return -1;
    h1[0] = lambda_deltat;
    h1[1] = kbar + (r2 / three) * hbari_np1 * lambda_deltat;
    h1[2] = h[2] + lambda_deltat * root23;
    h1[4] = hprime;
    double const1 = twthrd * hbark_np1 * lambda_deltat / mrts;
    double const2 =
        one - lambda_deltat * (twthrd * hbark_np1 + two * shear_mod) / mrts;
    for (int k = 0; k < 6; ++k) h1[5 + k] = lk * h[5 + k] + const1 * rtse[k];
    for (int k = 0; k < 6; ++k)
      devstr[k] = h1[5 + k] + b * h[5 + k] + const2 * rtse[k];
  } else {
    h1[0] = zero;
    h1[1] = kbar;
    if (lnelas) h1[1] = zero;
    h1[2] = h[2];
    h1[4] = hprime;
    for (int k = 0; k < 6; ++k) h1[5 + k] = alpha_n[k] * lk;
    for (int k = 0; k < 6; ++k) devstr[k] = rtse[k] + alpha_n[k];
  }
  /* ---- mm01_sig_final ---- */
  double sig_mean_np1 =
      eps_vol_n1 *
      (three * e * nu / ((one + nu) * (one - two * nu)) + two * shear_mod) /
      three;
  cgn1[0] = devstr[0] + sig_mean_np1;
  cgn1[1] = devstr[1] + sig_mean_np1;
  cgn1[2] = devstr[2] + sig_mean_np1;
  cgn1[3] = devstr[3];
  cgn1[4] = devstr[4];
  cgn1[5] = devstr[5];
  cgn1[6] = cgn[6] + half * (deps[0] * (cgn1[0] + cgn[0]) +
                             deps[1] * (cgn1[1] + cgn[1]) +
                             deps[2] * (cgn1[2] + cgn[2]) +
                             deps[3] * (cgn1[3] + cgn[3]) +
                             deps[4] * (cgn1[4] + cgn[4]) +
                             deps[5] * (cgn1[5] + cgn[5]));
  h1[3] = state_word(instat);
  /* ---- mm01_wrapup (the elastic strains are not kept) ---- */
  cgn1[7] = cgn[7];
  cgn1[8] = cgn[8];
  if (yield) {
    const double r2w = sqrt(2.0);
    const double* s = cgn;
    double sbar_n = sqrt((s[0] - s[1]) * (s[0] - s[1]) +
                         (s[1] - s[2]) * (s[1] - s[2]) +
                         (s[0] - s[2]) * (s[0] - s[2]) +
                         six * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])) /
                    r2w;
    s = cgn1;
    double sbar_n1 = sqrt((s[0] - s[1]) * (s[0] - s[1]) +
                          (s[1] - s[2]) * (s[1] - s[2]) +
                          (s[0] - s[2]) * (s[0] - s[2]) +
                          six * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])) /
                     r2w;
    double depbar = h1[2] - h[2];
    cgn1[7] = cgn[7] + half * (sbar_n1 + sbar_n) * depbar;
    cgn1[8] = cgn[8] + depbar;
  }
  /* ---- cnst1 with the updated history ---- */
  {
    const double root2c = 1.414213562;
    double kn1 = h1[1], hp = h1[4], ldt = h1[0];
    int y = state_of(h1[3]) == 1;
    for (int k = 0; k < 36; ++k) cep[k] = zero;
#define C(r, c) cep[6 * ((r)-1) + (c)-1]
    if (!y) {
      double c1 = (e / ((one + nu) * (one - two * nu)));
      double c2 = (one - nu) * c1;
      double c3 = ((one - two * nu) / two) * c1;
      double c4 = nu * c1;
      C(1, 1) = c2; C(2, 2) = c2; C(3, 3) = c2;
      C(4, 4) = c3; C(5, 5) = c3; C(6, 6) = c3;
      C(1, 2) = c4; C(1, 3) = c4; C(2, 1) = c4;
      C(3, 1) = c4; C(2, 3) = c4; C(3, 2) = c4;
    } else {
      const double* r = rtse; /* rtsg */
      double g = e / (two * (one + nu));
      double l = (e * nu) / ((one + nu) * (one - two * nu));
      double kk = (three * l + two * g) / three;
      double mrtsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] +
                     two * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
      double bb = (root2c * kn1 + (two / three) * (one - beta) * hp * ldt) /
                  sqrt(mrtsq);
      double gamma = one / (one + hp / (three * g));
      double gambar = gamma - one + bb;
      double gbar = g * bb;
      double albar = kk - two * gbar / three;
      double thbar = two * g * gambar;
      C(1, 1) = (albar + two * gbar - thbar * (r[0] * r[0]) / mrtsq);
      C(2, 2) = (albar + two * gbar - thbar * (r[1] * r[1]) / mrtsq);
      C(3, 3) = (albar + two * gbar - thbar * (r[2] * r[2]) / mrtsq);
      C(4, 4) = (gbar - thbar * (r[3] * r[3]) / mrtsq);
      C(5, 5) = (gbar - thbar * (r[4] * r[4]) / mrtsq);
      C(6, 6) = (gbar - thbar * (r[5] * r[5]) / mrtsq);
      C(2, 1) = (albar - thbar * r[0] * r[1] / mrtsq);
      C(3, 1) = (albar - thbar * r[0] * r[2] / mrtsq);
      C(4, 1) = -(thbar * r[0] * r[3] / mrtsq);
      C(5, 1) = -(thbar * r[0] * r[4] / mrtsq);
      C(6, 1) = -(thbar * r[0] * r[5] / mrtsq);
      C(3, 2) = (albar - thbar * r[2] * r[1] / mrtsq);
      C(4, 2) = -(thbar * r[1] * r[3] / mrtsq);
      C(5, 2) = -(thbar * r[1] * r[4] / mrtsq);
      C(6, 2) = -(thbar * r[1] * r[5] / mrtsq);
      C(4, 3) = -(thbar * r[2] * r[3] / mrtsq);
      C(5, 3) = -(thbar * r[2] * r[4] / mrtsq);
      C(6, 3) = -(thbar * r[2] * r[5] / mrtsq);
      C(5, 4) = -(thbar * r[3] * r[4] / mrtsq);
      C(6, 4) = -(thbar * r[3] * r[5] / mrtsq);
      C(6, 5) = -(thbar * r[4] * r[5] / mrtsq);
      for (int a = 1; a <= 6; ++a)
        for (int c = a + 1; c <= 6; ++c) C(a, c) = C(c, a);
    }
#undef C
  }
  return 0;
}
