// w3d_geonl.h -- device helpers of the finite-strain (geonl) path, for one
// Gauss point in registers.
//
// Reference (src/): gtmat1.f:522 tcomp1, :601 fcomp1, :691 rtcmp1, :740
// getrm1, :961 irscp1, :1031 ivcmp1, :1134 evcmp1_new (eigenvalues of a
// symmetric 3x3 by the trigonometric solution of its characteristic cubic,
// the published closed form of J. Kopp, Int. J. Mod. Phys. C 19 (2008)
// 523); qmply1.f:15; gptns1.f:1735 ctran1 and :1905 ctran1_qbar.
//
// 3x3 matrices are row-major (a[3 r + c]); symmetric 3x3 tensors in
// "upper" form are (11, 12, 22, 13, 23, 33); vectors of stress and strain
// are (xx, yy, zz, xy, yz, xz); 6x6 matrices row-major.
#pragma once

#include <hip/hip_runtime.h>

namespace w3d {

// F = I + grad u from the n = 0 inverse Jacobian ga (row-major), the
// parametric derivatives of the point and the element displacements
// u = ue + cf due (x-block, y-block, z-block of 3 NN words; rstgp1_a);
// returns det F.
template <int NN>
__device__ inline double geonl_f(const double* ga, const double* nx,
                                 const double* ny, const double* nz,
                                 const double* ue, const double* due,
                                 double cf, double f[9]) {
  double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int j = 0; j < NN; ++j) {
    const double t1 = ga[0] * nx[j] + ga[1] * ny[j] + ga[2] * nz[j];
    const double t2 = ga[3] * nx[j] + ga[4] * ny[j] + ga[5] * nz[j];
    const double t3 = ga[6] * nx[j] + ga[7] * ny[j] + ga[8] * nz[j];
    const double a = ue[j] + cf * due[j], b = ue[NN + j] + cf * due[NN + j],
                 c = ue[2 * NN + j] + cf * due[2 * NN + j];
    t[0] = t[0] + t1 * a;
    t[1] = t[1] + t1 * b;
    t[2] = t[2] + t1 * c;
    t[3] = t[3] + t2 * a;
    t[4] = t[4] + t2 * b;
    t[5] = t[5] + t2 * c;
    t[6] = t[6] + t3 * a;
    t[7] = t[7] + t3 * b;
    t[8] = t[8] + t3 * c;
  }
  // F(r, c) = d x_r / d X_c: row r gathers the gradients of u_r
  f[0] = t[0] + 1.0; f[1] = t[3]; f[2] = t[6];
  f[3] = t[1]; f[4] = t[4] + 1.0; f[5] = t[7];
  f[6] = t[2]; f[7] = t[5]; f[8] = t[8] + 1.0;
  const double f1 = f[4] * f[8] - f[5] * f[7];
  const double f2 = f[3] * f[8] - f[5] * f[6];
  const double f3 = f[3] * f[7] - f[4] * f[6];
  return f[0] * f1 - f[1] * f2 + f[2] * f3;
}

// Rotation R of the polar decomposition F = R U, with U^-1 from the
// principal invariants of U: U^-1 = a2 (b2 I + c2 C + d2 C^2), C = F^T F.
__device__ inline void geonl_rotation(const double f[9], double r[9]) {
  double c[6], cc[6];
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
  // eigenvalues of C from its characteristic cubic (trigonometric form)
  const double third = 1.0 / 3.0, iroot3 = 1.0 / sqrt(3.0);
  const double m11 = c[0], m12 = c[1], m13 = c[3], m22 = c[2], m23 = c[4],
               m33 = c[5];
  const double de = m12 * m23, dd = m12 * m12, ee = m23 * m23, ff = m13 * m13;
  const double m = m11 + m22 + m33;
  const double c1 = (m11 * m22 + m11 * m33 + m22 * m33) - (dd + ee + ff);
  const double c0 = m33 * dd + m11 * ee + m22 * ff - m11 * m22 * m33 -
                    2.0 * m13 * de;
  const double p = m * m - 3.0 * c1;
  const double q = m * (p - 1.5 * c1) - 13.5 * c0;
  const double sqrtp = sqrt(fabs(p));
  double phi = 27.0 * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0));
  phi = third * atan2(sqrt(fabs(phi)), q);
  const double cphi = sqrtp * cos(phi);
  const double sphi = iroot3 * sqrtp * sin(phi);
  double e2 = third * (m - cphi);
  const double e3 = e2 + sphi;
  const double e1 = e2 + cphi;
  e2 = e2 - sphi;
  // ascending, as the invariant sums below take them
  double l1 = e1, l2 = e2, l3 = e3, s;
  if (l2 < l1) { s = l1; l1 = l2; l2 = s; }
  if (l3 < l1) { s = l1; l1 = l3; l3 = s; }
  if (l3 < l2) { s = l2; l2 = l3; l3 = s; }
  l1 = sqrt(l1);
  l2 = sqrt(l2);
  l3 = sqrt(l3);
  const double iu = l1 + l2 + l3;
  const double iiu = l1 * l2 + l2 * l3 + l1 * l3;
  const double iiiu = l1 * l2 * l3;
  const double a2 = 1.0 / (iiiu * (iu * iiu - iiiu));
  const double b2 = iu * iiu * iiu - iiiu * (iu * iu + iiu);
  const double c2 = -iiiu - iu * (iu * iu - 2.0 * iiu);
  const double d2 = iu;
  double ui[6];
  ui[0] = a2 * (b2 + c2 * c[0] + d2 * cc[0]);
  ui[1] = a2 * (c2 * c[1] + d2 * cc[1]);
  ui[2] = a2 * (b2 + c2 * c[2] + d2 * cc[2]);
  ui[3] = a2 * (c2 * c[3] + d2 * cc[3]);
  ui[4] = a2 * (c2 * c[4] + d2 * cc[4]);
  ui[5] = a2 * (b2 + c2 * c[5] + d2 * cc[5]);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r[3 * a + 0] = f[3 * a] * ui[0] + f[3 * a + 1] * ui[1] + f[3 * a + 2] * ui[3];
    r[3 * a + 1] = f[3 * a] * ui[1] + f[3 * a + 1] * ui[2] + f[3 * a + 2] * ui[4];
    r[3 * a + 2] = f[3 * a] * ui[3] + f[3 * a + 1] * ui[4] + f[3 * a + 2] * ui[5];
  }
}

// 6x6 rotation of symmetric tensors in vector form (getrm1):
// opt 1: {d} = q {D} with [d] = R^T [D] R, engineering shear;
// opt 2: {T} = q {t} with [T] = R [t] R^T.
__device__ inline void geonl_q(const double* r9, int opt, double* q) {
#define R_(a, b) r9[3 * ((a)-1) + ((b)-1)]
#define Q_(a, b) q[6 * ((a)-1) + ((b)-1)]
  if (opt == 1) {
    Q_(1, 1) = R_(1, 1) * R_(1, 1); Q_(1, 2) = R_(2, 1) * R_(2, 1);
    Q_(1, 3) = R_(3, 1) * R_(3, 1); Q_(1, 4) = R_(1, 1) * R_(2, 1);
    Q_(1, 5) = R_(3, 1) * R_(2, 1); Q_(1, 6) = R_(1, 1) * R_(3, 1);
    Q_(2, 1) = R_(1, 2) * R_(1, 2); Q_(2, 2) = R_(2, 2) * R_(2, 2);
    Q_(2, 3) = R_(3, 2) * R_(3, 2); Q_(2, 4) = R_(1, 2) * R_(2, 2);
    Q_(2, 5) = R_(3, 2) * R_(2, 2); Q_(2, 6) = R_(1, 2) * R_(3, 2);
    Q_(3, 1) = R_(1, 3) * R_(1, 3); Q_(3, 2) = R_(2, 3) * R_(2, 3);
    Q_(3, 3) = R_(3, 3) * R_(3, 3); Q_(3, 4) = R_(1, 3) * R_(2, 3);
    Q_(3, 5) = R_(3, 3) * R_(2, 3); Q_(3, 6) = R_(1, 3) * R_(3, 3);
    Q_(4, 1) = 2.0 * R_(1, 1) * R_(1, 2); Q_(4, 2) = 2.0 * R_(2, 1) * R_(2, 2);
    Q_(4, 3) = 2.0 * R_(3, 1) * R_(3, 2);
    Q_(4, 4) = R_(1, 1) * R_(2, 2) + R_(1, 2) * R_(2, 1);
    Q_(4, 5) = R_(2, 1) * R_(3, 2) + R_(3, 1) * R_(2, 2);
    Q_(4, 6) = R_(1, 1) * R_(3, 2) + R_(3, 1) * R_(1, 2);
    Q_(5, 1) = 2.0 * R_(1, 2) * R_(1, 3); Q_(5, 2) = 2.0 * R_(2, 3) * R_(2, 2);
    Q_(5, 3) = 2.0 * R_(3, 2) * R_(3, 3);
    Q_(5, 4) = R_(1, 2) * R_(2, 3) + R_(2, 2) * R_(1, 3);
    Q_(5, 5) = R_(2, 2) * R_(3, 3) + R_(2, 3) * R_(3, 2);
    Q_(5, 6) = R_(1, 2) * R_(3, 3) + R_(3, 2) * R_(1, 3);
    Q_(6, 1) = 2.0 * R_(1, 1) * R_(1, 3); Q_(6, 2) = 2.0 * R_(2, 1) * R_(2, 3);
    Q_(6, 3) = 2.0 * R_(3, 1) * R_(3, 3);
    Q_(6, 4) = R_(1, 1) * R_(2, 3) + R_(2, 1) * R_(1, 3);
    Q_(6, 5) = R_(2, 1) * R_(3, 3) + R_(3, 1) * R_(2, 3);
    Q_(6, 6) = R_(1, 1) * R_(3, 3) + R_(1, 3) * R_(3, 1);
  } else {
    Q_(1, 1) = R_(1, 1) * R_(1, 1); Q_(1, 2) = R_(1, 2) * R_(1, 2);
    Q_(1, 3) = R_(1, 3) * R_(1, 3); Q_(1, 4) = 2.0 * R_(1, 1) * R_(1, 2);
    Q_(1, 5) = 2.0 * R_(1, 3) * R_(1, 2); Q_(1, 6) = 2.0 * R_(1, 1) * R_(1, 3);
    Q_(2, 1) = R_(2, 1) * R_(2, 1); Q_(2, 2) = R_(2, 2) * R_(2, 2);
    Q_(2, 3) = R_(2, 3) * R_(2, 3); Q_(2, 4) = 2.0 * R_(2, 1) * R_(2, 2);
    Q_(2, 5) = 2.0 * R_(2, 3) * R_(2, 2); Q_(2, 6) = 2.0 * R_(2, 1) * R_(2, 3);
    Q_(3, 1) = R_(3, 1) * R_(3, 1); Q_(3, 2) = R_(3, 2) * R_(3, 2);
    Q_(3, 3) = R_(3, 3) * R_(3, 3); Q_(3, 4) = 2.0 * R_(3, 1) * R_(3, 2);
    Q_(3, 5) = 2.0 * R_(3, 3) * R_(3, 2); Q_(3, 6) = 2.0 * R_(3, 1) * R_(3, 3);
    Q_(4, 1) = R_(1, 1) * R_(2, 1); Q_(4, 2) = R_(1, 2) * R_(2, 2);
    Q_(4, 3) = R_(1, 3) * R_(2, 3);
    Q_(4, 4) = R_(1, 1) * R_(2, 2) + R_(2, 1) * R_(1, 2);
    Q_(4, 5) = R_(1, 2) * R_(2, 3) + R_(1, 3) * R_(2, 2);
    Q_(4, 6) = R_(1, 1) * R_(2, 3) + R_(1, 3) * R_(2, 1);
    Q_(5, 1) = R_(2, 1) * R_(3, 1); Q_(5, 2) = R_(3, 2) * R_(2, 2);
    Q_(5, 3) = R_(2, 3) * R_(3, 3);
    Q_(5, 4) = R_(2, 1) * R_(3, 2) + R_(2, 2) * R_(3, 1);
    Q_(5, 5) = R_(2, 2) * R_(3, 3) + R_(3, 2) * R_(2, 3);
    Q_(5, 6) = R_(2, 1) * R_(3, 3) + R_(2, 3) * R_(3, 1);
    Q_(6, 1) = R_(1, 1) * R_(3, 1); Q_(6, 2) = R_(1, 2) * R_(3, 2);
    Q_(6, 3) = R_(1, 3) * R_(3, 3);
    Q_(6, 4) = R_(1, 1) * R_(3, 2) + R_(1, 2) * R_(3, 1);
    Q_(6, 5) = R_(1, 2) * R_(3, 3) + R_(1, 3) * R_(3, 2);
    Q_(6, 6) = R_(1, 1) * R_(3, 3) + R_(3, 1) * R_(1, 3);
  }
#undef R_
#undef Q_
}

// out = q in (qmply1)
__device__ inline void geonl_qmul(const double* q, const double* in,
                                  double* out) {
#pragma unroll
  for (int j = 0; j < 6; ++j)
    out[j] = q[6 * j + 0] * in[0] + q[6 * j + 1] * in[1] + q[6 * j + 2] * in[2] +
             q[6 * j + 3] * in[3] + q[6 * j + 4] * in[4] + q[6 * j + 5] * in[5];
}

// ctran1 + ctran1_qbar on a full row-major 6x6 D: D* = q D q^T, then the
// initial-stress matrix of the Cauchy stress cs times wf = w |J| taken off
// (its shear-shear couplings as the reference has them).
__device__ inline void geonl_ctran(double* d, const double* q,
                                   const double* cs, double wf) {
  double tc[36];
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int m = 0; m < 6; ++m)
      tc[6 * j + m] = (q[6 * j + 0] * d[m] + q[6 * j + 1] * d[6 + m] +
                       q[6 * j + 2] * d[12 + m] + q[6 * j + 3] * d[18 + m] +
                       q[6 * j + 4] * d[24 + m] + q[6 * j + 5] * d[30 + m]);
#pragma unroll
  for (int j = 0; j < 6; ++j)
#pragma unroll
    for (int m = 0; m < 6; ++m)
      d[6 * j + m] = tc[6 * j + 0] * q[6 * m + 0] + tc[6 * j + 1] * q[6 * m + 1] +
                     tc[6 * j + 2] * q[6 * m + 2] + tc[6 * j + 3] * q[6 * m + 3] +
                     tc[6 * j + 4] * q[6 * m + 4] + tc[6 * j + 5] * q[6 * m + 5];
  const double halfw = 0.5 * wf;
  const double xx = cs[0], yy = cs[1], zz = cs[2], xy = cs[3], yz = cs[4],
               xz = cs[5];
  d[0] = d[0] - 2.0 * xx * wf;
  d[7] = d[7] - 2.0 * yy * wf;
  d[14] = d[14] - 2.0 * zz * wf;
  d[18] = d[18] - xy * wf;  // (4,1)
  d[30] = d[30] - xz * wf;  // (6,1)
  d[19] = d[19] - xy * wf;  // (4,2)
  d[25] = d[25] - yz * wf;  // (5,2)
  d[26] = d[26] - yz * wf;  // (5,3)
  d[32] = d[32] - xz * wf;  // (6,3)
  d[21] = d[21] - halfw * (xx + yy);
  d[28] = d[28] - halfw * (yy + zz);
  d[35] = d[35] - halfw * (xx + zz);
  d[27] = d[27] - wf * xz;  // (5,4)
  d[33] = d[33] - wf * yz;  // (6,4)
  d[34] = d[34] - wf * xy;  // (6,5)
  d[3] = d[18];
  d[5] = d[30];
  d[9] = d[19];
  d[10] = d[25];
  d[16] = d[26];
  d[17] = d[32];
  d[22] = d[27];
  d[23] = d[33];
  d[29] = d[34];
}

}  // namespace w3d
