/* orc_geonl.c -- CPU restatement (test infrastructure: the oracle) of the
 * finite-strain (geonl) kinematics of WARP3D's element-block path.
 *
 * Reference (src/):
 *   rstgp1.f:19 rstgp1, :5168 rstgp1_a, :5189 rstgp1_b
 *   gtmat1.f:17 gtmat1, :522 tcomp1, :601 fcomp1, :691 rtcmp1, :740 getrm1,
 *     :961 irscp1, :1031 ivcmp1, :1134 evcmp1_new (eigenvalues of a
 *     symmetric 3x3 by Cardano's formula; restated from the published
 *     closed form, J. Kopp, "Efficient numerical diagonalization of
 *     hermitian 3x3 matrices", Int. J. Mod. Phys. C 19 (2008) 523)
 *   qmply1.f:15 qmply1
 *   gptns1.f:1735 ctran1 (+ ctran1_qbar :1905), kg1.f:16 kgstiff
 *
 * Layouts: F, R and the Jacobian inverse are row-major 3x3 here
 * (F[3 r + c] = F(r+1, c+1)); the reference's 9-word rotation records
 * (rot_blk_n1(i, 1:9, gp)) are column-major and converted where they are
 * stored.  Symmetric 3x3 tensors in the "upper triangular form" of
 * ivcmp1/irscp1 are (11, 12, 22, 13, 23, 33); stress and strain vectors
 * are (xx, yy, zz, xy, yz, xz).
 */
#include <math.h>
#include <string.h>

#include "w3d_oracle.h"

/* tcomp1: displacement gradient theta(1..9) at a point: with
   g_j = gama dN_j (the spatial gradient of node j's shape function),
   theta = sum_j (g_j,x ux, g_j,x uy, g_j,x uz, g_j,y ux, ..., g_j,z uz) */
void orc_tcomp1(int nnode, const double gama[9], const double* nxi,
                const double* neta, const double* nzeta, const double* ue,
                double theta[9]) {
  for (int k = 0; k < 9; ++k) theta[k] = 0.0;
  for (int j = 0; j < nnode; ++j) {
    double t1 = gama[0] * nxi[j] + gama[1] * neta[j] + gama[2] * nzeta[j];
    double t2 = gama[3] * nxi[j] + gama[4] * neta[j] + gama[5] * nzeta[j];
    double t3 = gama[6] * nxi[j] + gama[7] * neta[j] + gama[8] * nzeta[j];
    double ux = ue[j], uy = ue[nnode + j], uz = ue[2 * nnode + j];
    theta[0] = theta[0] + t1 * ux;
    theta[1] = theta[1] + t1 * uy;
    theta[2] = theta[2] + t1 * uz;
    theta[3] = theta[3] + t2 * ux;
    theta[4] = theta[4] + t2 * uy;
    theta[5] = theta[5] + t2 * uz;
    theta[6] = theta[6] + t3 * ux;
    theta[7] = theta[7] + t3 * uy;
    theta[8] = theta[8] + t3 * uz;
  }
}

/* fcomp1: F = I + grad u and det F; returns 1 when det F <= 1e-20 */
int orc_fcomp1(const double theta[9], double f[9], double* df) {
  f[0] = theta[0] + 1.0; f[1] = theta[3]; f[2] = theta[6];
  f[3] = theta[1]; f[4] = theta[4] + 1.0; f[5] = theta[7];
  f[6] = theta[2]; f[7] = theta[5]; f[8] = theta[8] + 1.0;
  double f1 = f[4] * f[8] - f[5] * f[7];
  double f2 = f[3] * f[8] - f[5] * f[6];
  double f3 = f[3] * f[7] - f[4] * f[6];
  *df = f[0] * f1 - f[1] * f2 + f[2] * f3;
  return *df <= 1.0e-20 ? 1 : 0;
}

/* evcmp1_new: eigenvalues (ascending) of the symmetric 3x3 c = (11, 12,
   22, 13, 23, 33) from the characteristic cubic in trigonometric form */
void orc_evcmp1(const double c[6], double ev[3]) {
  const double third = 1.0 / 3.0, iroot3 = 1.0 / sqrt(3.0);
  double m11 = c[0], m12 = c[1], m13 = c[3], m22 = c[2], m23 = c[4],
         m33 = c[5];
  double de = m12 * m23;
  double dd = m12 * m12;
  double ee = m23 * m23;
  double ff = m13 * m13;
  double m = m11 + m22 + m33;
  double c1 = (m11 * m22 + m11 * m33 + m22 * m33) - (dd + ee + ff);
  double c0 = m33 * dd + m11 * ee + m22 * ff - m11 * m22 * m33 -
              2.0 * m13 * de;
  double p = m * m - 3.0 * c1;
  double q = m * (p - 1.5 * c1) - 13.5 * c0;
  double sqrtp = sqrt(fabs(p));
  double phi = 27.0 * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0));
  phi = third * atan2(sqrt(fabs(phi)), q);
  double cphi = sqrtp * cos(phi);
  double sphi = iroot3 * sqrtp * sin(phi);
  double e2 = third * (m - cphi);
  double e3 = e2 + sphi;
  double e1 = e2 + cphi;
  e2 = e2 - sphi;
  double s;
  if (e2 < e1) { s = e1; e1 = e2; e2 = s; }
  if (e3 < e1) { s = e1; e1 = e3; e3 = s; }
  if (e3 < e2) { s = e2; e2 = e3; e3 = s; }
  ev[0] = e1;
  ev[1] = e2;
  ev[2] = e3;
}

/* ivcmp1 + irscp1: the inverse right stretch U^-1 = a2 (b2 I + c2 C +
   d2 C^2) from the invariants of U (C = F^T F), upper form (11, 12, 22,
   13, 23, 33) */
void orc_irscp1(const double f[9], double ui[6]) {
  double c[6], cc[6], ev[3];
  c[0] = f[0] * f[0] + f[3] * f[3] + f[6] * f[6];
  c[1] = f[0] * f[1] + f[3] * f[4] + f[6] * f[7];
  c[2] = f[1] * f[1] + f[4] * f[4] + f[7] * f[7];
  c[3] = f[0] * f[2] + f[3] * f[5] + f[6] * f[8];
  c[4] = f[1] * f[2] + f[4] * f[5] + f[7] * f[8];
  c[5] = f[2] * f[2] + f[5] * f[5] + f[8] * f[8];
  cc[0] = c[0] * c[0] + c[1] * c[1] + c[3] * c[3];
  cc[1] = c[0] * c[1] + c[1] * c[2] + c[3] * c[4];
  cc[2] = c[1] * c[1] + c[2] * c[2] + c[4] * c[4];
  cc[3] = c[0] * c[3] + c[1] * c[4] + c[3] * c[5];
  cc[4] = c[1] * c[3] + c[2] * c[4] + c[4] * c[5];
  cc[5] = c[3] * c[3] + c[4] * c[4] + c[5] * c[5];
  orc_evcmp1(c, ev);
  ev[0] = sqrt(ev[0]);
  ev[1] = sqrt(ev[1]);
  ev[2] = sqrt(ev[2]);
  double iu = ev[0] + ev[1] + ev[2];
  double iiu = ev[0] * ev[1] + ev[1] * ev[2] + ev[0] * ev[2];
  double iiiu = ev[0] * ev[1] * ev[2];
  double a2 = 1.0 / (iiiu * (iu * iiu - iiiu));
  double b2 = iu * iiu * iiu - iiiu * (iu * iu + iiu);
  double c2 = -iiiu - iu * (iu * iu - 2.0 * iiu);
  double d2 = iu;
  ui[0] = a2 * (b2 + c2 * c[0] + d2 * cc[0]);
  ui[1] = a2 * (c2 * c[1] + d2 * cc[1]);
  ui[2] = a2 * (b2 + c2 * c[2] + d2 * cc[2]);
  ui[3] = a2 * (c2 * c[3] + d2 * cc[3]);
  ui[4] = a2 * (c2 * c[4] + d2 * cc[4]);
  ui[5] = a2 * (b2 + c2 * c[5] + d2 * cc[5]);
}

/* rtcmp1: R = F U^-1 */
void orc_rtcmp1(const double f[9], double r[9]) {
  double ui[6];
  orc_irscp1(f, ui);
  for (int a = 0; a < 3; ++a) {
    const double* fr = f + 3 * a;
    r[3 * a + 0] = fr[0] * ui[0] + fr[1] * ui[1] + fr[2] * ui[3];
    r[3 * a + 1] = fr[0] * ui[1] + fr[1] * ui[2] + fr[2] * ui[4];
    r[3 * a + 2] = fr[0] * ui[3] + fr[1] * ui[4] + fr[2] * ui[5];
  }
}

/* getrm1: 6x6 transformation q (row-major q[6 j + k]) from R.
   opt 1: unrotated rate of deformation {d} = q {D} ([d] = R^T [D] R,
          engineering shear);
   opt 2: Cauchy stress {T} = q {t} ([T] = R [t] R^T). */
void orc_getrm1(const double r9[9], int opt, double q[36]) {
#define R(a, b) r9[3 * ((a)-1) + ((b)-1)]
#define Q(a, b) q[6 * ((a)-1) + ((b)-1)]
  if (opt == 1) {
    Q(1, 1) = R(1, 1) * R(1, 1); Q(1, 2) = R(2, 1) * R(2, 1);
    Q(1, 3) = R(3, 1) * R(3, 1); Q(1, 4) = R(1, 1) * R(2, 1);
    Q(1, 5) = R(3, 1) * R(2, 1); Q(1, 6) = R(1, 1) * R(3, 1);
    Q(2, 1) = R(1, 2) * R(1, 2); Q(2, 2) = R(2, 2) * R(2, 2);
    Q(2, 3) = R(3, 2) * R(3, 2); Q(2, 4) = R(1, 2) * R(2, 2);
    Q(2, 5) = R(3, 2) * R(2, 2); Q(2, 6) = R(1, 2) * R(3, 2);
    Q(3, 1) = R(1, 3) * R(1, 3); Q(3, 2) = R(2, 3) * R(2, 3);
    Q(3, 3) = R(3, 3) * R(3, 3); Q(3, 4) = R(1, 3) * R(2, 3);
    Q(3, 5) = R(3, 3) * R(2, 3); Q(3, 6) = R(1, 3) * R(3, 3);
    Q(4, 1) = 2.0 * R(1, 1) * R(1, 2); Q(4, 2) = 2.0 * R(2, 1) * R(2, 2);
    Q(4, 3) = 2.0 * R(3, 1) * R(3, 2);
    Q(4, 4) = R(1, 1) * R(2, 2) + R(1, 2) * R(2, 1);
    Q(4, 5) = R(2, 1) * R(3, 2) + R(3, 1) * R(2, 2);
    Q(4, 6) = R(1, 1) * R(3, 2) + R(3, 1) * R(1, 2);
    Q(5, 1) = 2.0 * R(1, 2) * R(1, 3); Q(5, 2) = 2.0 * R(2, 3) * R(2, 2);
    Q(5, 3) = 2.0 * R(3, 2) * R(3, 3);
    Q(5, 4) = R(1, 2) * R(2, 3) + R(2, 2) * R(1, 3);
    Q(5, 5) = R(2, 2) * R(3, 3) + R(2, 3) * R(3, 2);
    Q(5, 6) = R(1, 2) * R(3, 3) + R(3, 2) * R(1, 3);
    Q(6, 1) = 2.0 * R(1, 1) * R(1, 3); Q(6, 2) = 2.0 * R(2, 1) * R(2, 3);
    Q(6, 3) = 2.0 * R(3, 1) * R(3, 3);
    Q(6, 4) = R(1, 1) * R(2, 3) + R(2, 1) * R(1, 3);
    Q(6, 5) = R(2, 1) * R(3, 3) + R(3, 1) * R(2, 3);
    Q(6, 6) = R(1, 1) * R(3, 3) + R(1, 3) * R(3, 1);
  } else {
    Q(1, 1) = R(1, 1) * R(1, 1); Q(1, 2) = R(1, 2) * R(1, 2);
    Q(1, 3) = R(1, 3) * R(1, 3); Q(1, 4) = 2.0 * R(1, 1) * R(1, 2);
    Q(1, 5) = 2.0 * R(1, 3) * R(1, 2); Q(1, 6) = 2.0 * R(1, 1) * R(1, 3);
    Q(2, 1) = R(2, 1) * R(2, 1); Q(2, 2) = R(2, 2) * R(2, 2);
    Q(2, 3) = R(2, 3) * R(2, 3); Q(2, 4) = 2.0 * R(2, 1) * R(2, 2);
    Q(2, 5) = 2.0 * R(2, 3) * R(2, 2); Q(2, 6) = 2.0 * R(2, 1) * R(2, 3);
    Q(3, 1) = R(3, 1) * R(3, 1); Q(3, 2) = R(3, 2) * R(3, 2);
    Q(3, 3) = R(3, 3) * R(3, 3); Q(3, 4) = 2.0 * R(3, 1) * R(3, 2);
    Q(3, 5) = 2.0 * R(3, 3) * R(3, 2); Q(3, 6) = 2.0 * R(3, 1) * R(3, 3);
    Q(4, 1) = R(1, 1) * R(2, 1); Q(4, 2) = R(1, 2) * R(2, 2);
    Q(4, 3) = R(1, 3) * R(2, 3);
    Q(4, 4) = R(1, 1) * R(2, 2) + R(2, 1) * R(1, 2);
    Q(4, 5) = R(1, 2) * R(2, 3) + R(1, 3) * R(2, 2);
    Q(4, 6) = R(1, 1) * R(2, 3) + R(1, 3) * R(2, 1);
    Q(5, 1) = R(2, 1) * R(3, 1); Q(5, 2) = R(3, 2) * R(2, 2);
    Q(5, 3) = R(2, 3) * R(3, 3);
    Q(5, 4) = R(2, 1) * R(3, 2) + R(2, 2) * R(3, 1);
    Q(5, 5) = R(2, 2) * R(3, 3) + R(3, 2) * R(2, 3);
    Q(5, 6) = R(2, 1) * R(3, 3) + R(2, 3) * R(3, 1);
    Q(6, 1) = R(1, 1) * R(3, 1); Q(6, 2) = R(1, 2) * R(3, 2);
    Q(6, 3) = R(1, 3) * R(3, 3);
    Q(6, 4) = R(1, 1) * R(3, 2) + R(1, 2) * R(3, 1);
    Q(6, 5) = R(1, 2) * R(3, 3) + R(1, 3) * R(3, 2);
    Q(6, 6) = R(1, 1) * R(3, 3) + R(3, 1) * R(1, 3);
  }
#undef R
#undef Q
}

/* qmply1: m2 = q m1 (6x6 times 6) */
void orc_qmply1(const double q[36], const double m1[6], double m2[6]) {
  for (int j = 0; j < 6; ++j)
    m2[j] = q[6 * j + 0] * m1[0] + q[6 * j + 1] * m1[1] +
            q[6 * j + 2] * m1[2] + q[6 * j + 3] * m1[3] +
            q[6 * j + 4] * m1[4] + q[6 * j + 5] * m1[5];
}

/* ctran1: D* = q D q^T (D row-major, q = getrm1 opt 2 of R n+1), then
   with qbar the initial-stress correction of the Cauchy stress cs
   (scaled by w |J|, like the D it modifies) */
void orc_ctran1(double d[36], const double q[36], const double cs[6],
                int qbar, double wdj) {
  double tc[36];
  for (int j = 0; j < 6; ++j)
    for (int m = 0; m < 6; ++m)
      tc[6 * j + m] = (q[6 * j + 0] * d[0 * 6 + m] + q[6 * j + 1] * d[1 * 6 + m] +
                       q[6 * j + 2] * d[2 * 6 + m] + q[6 * j + 3] * d[3 * 6 + m] +
                       q[6 * j + 4] * d[4 * 6 + m] + q[6 * j + 5] * d[5 * 6 + m]);
  for (int j = 0; j < 6; ++j)
    for (int m = 0; m < 6; ++m)
      d[6 * j + m] = tc[6 * j + 0] * q[6 * m + 0] + tc[6 * j + 1] * q[6 * m + 1] +
                     tc[6 * j + 2] * q[6 * m + 2] + tc[6 * j + 3] * q[6 * m + 3] +
                     tc[6 * j + 4] * q[6 * m + 4] + tc[6 * j + 5] * q[6 * m + 5];
  if (!qbar) return;
#define D(a, b) d[6 * ((a)-1) + ((b)-1)]
  const double wf = wdj, halfw = 0.5 * wf;
  const double xx = cs[0], yy = cs[1], zz = cs[2], xy = cs[3], yz = cs[4],
               xz = cs[5];
  D(1, 1) = D(1, 1) - 2.0 * xx * wf;
  D(2, 2) = D(2, 2) - 2.0 * yy * wf;
  D(3, 3) = D(3, 3) - 2.0 * zz * wf;
  D(4, 1) = D(4, 1) - xy * wf;
  D(6, 1) = D(6, 1) - xz * wf;
  D(4, 2) = D(4, 2) - xy * wf;
  D(5, 2) = D(5, 2) - yz * wf;
  D(5, 3) = D(5, 3) - yz * wf;
  D(6, 3) = D(6, 3) - xz * wf;
  D(4, 4) = D(4, 4) - halfw * (xx + yy);
  D(5, 5) = D(5, 5) - halfw * (yy + zz);
  D(6, 6) = D(6, 6) - halfw * (xx + zz);
  D(5, 4) = D(5, 4) - wf * xz;
  D(6, 4) = D(6, 4) - wf * yz;
  D(6, 5) = D(6, 5) - wf * xy;
  D(1, 4) = D(4, 1);
  D(1, 6) = D(6, 1);
  D(2, 4) = D(4, 2);
  D(2, 5) = D(5, 2);
  D(3, 5) = D(5, 3);
  D(3, 6) = D(6, 3);
  D(4, 5) = D(5, 4);
  D(4, 6) = D(6, 4);
  D(5, 6) = D(6, 5);
#undef D
}

/* kgstiff: the geometric stiffness gtg(a, b) = (g_a^T S g_b) |J| w of
   node pair (a, b), added to the three component blocks of Ke (packed
   upper column-major, pos = cp(q) + p).  g are the spatial gradients of
   the shape functions at the point, or with B-bar the element's
   volume-averaged gradients vol (as the reference takes them). */
void orc_kgstiff(int nnode, const double gama[9], const double* nxi,
                 const double* neta, const double* nzeta, const double* vol,
                 const double sig[6], double dj, double w, double* ek) {
  double gx[32], gy[32], gz[32];
  for (int j = 0; j < nnode; ++j) {
    if (vol) {
      gx[j] = vol[3 * j];
      gy[j] = vol[3 * j + 1];
      gz[j] = vol[3 * j + 2];
    } else {
      gx[j] = gama[0] * nxi[j] + gama[1] * neta[j] + gama[2] * nzeta[j];
      gy[j] = gama[3] * nxi[j] + gama[4] * neta[j] + gama[5] * nzeta[j];
      gz[j] = gama[6] * nxi[j] + gama[7] * neta[j] + gama[8] * nzeta[j];
    }
  }
  for (int b = 0; b < nnode; ++b)
    for (int a = 0; a <= b; ++a) {
      double g = (gx[a] * gx[b] * sig[0] + gy[a] * gy[b] * sig[1] +
                  gz[a] * gz[b] * sig[2] +
                  (gx[a] * gy[b] + gx[b] * gy[a]) * sig[3] +
                  (gy[a] * gz[b] + gy[b] * gz[a]) * sig[4] +
                  (gx[a] * gz[b] + gx[b] * gz[a]) * sig[5]) *
                 dj * w;
      for (int c = 0; c < 3; ++c) {
        int p = c * nnode + a, q = c * nnode + b;
        int64_t pos = (int64_t)q * (q + 1) / 2 + p;
        ek[pos] = ek[pos] + g;
      }
    }
}
