/*
 * orc_elements.c -- oracle: Gauss rules, shape-function derivatives,
 * coordinate Jacobians, strain-displacement matrices.
 *
 * TEST INFRASTRUCTURE ONLY (see w3d_oracle.h).  Each routine restates the
 * reference routine named in its comment, keeping the reference's constants
 * (including truncated ones) and its point and node numbering.
 */
#include <math.h>
#include <string.h>
#include "w3d_oracle.h"

int orc_nnode(int etype) {
  switch (etype) {
    case W3D_Q3DISOP: return 20;
    case W3D_L3DISOP: return 8;
    case W3D_TET10: return 10;
    default: return 0;
  }
}

/* number of points per rule: elprp.f:510-521 (hex20), :741-748 (hex8),
   :1689-1698 (tet10) */
int orc_ngp(int etype, int order) {
  if (etype == W3D_L3DISOP) return order == 1 ? 8 : 0;
  if (etype == W3D_Q3DISOP) {
    switch (order) {
      case 1: return 27;
      case 2: return 9;
      case 8: return 8;
      case 9: return 14;
      default: return 0;
    }
  }
  if (etype == W3D_TET10) {
    if (order == 1 || order == 4 || order == 5) return order;
    return 0;
  }
  return 0;
}

/* quad1, getgpts.f:177 -- 20-node brick rules.  27-point: eta fastest, then
   xi, then zeta (getgpts.f:221-355). */
static void quad1(int order, int gpn, double* xi, double* eta, double* zeta,
                  double* w) {
  if (order == 1) {
    const double l1 = 0.774596669241483, l2 = 0.0;
    const double w1 = 0.1714677640603571, w2 = 0.2743484224965712,
                 w3 = 0.4389574759945135, w4 = 0.7023319615912210;
    const double lv[3] = {-l1, l2, l1};
    int k = gpn - 1;
    int ie = k % 3, ix = (k / 3) % 3, iz = k / 9;
    *xi = lv[ix];
    *eta = lv[ie];
    *zeta = lv[iz];
    int nmid = (ix == 1) + (ie == 1) + (iz == 1);
    *w = nmid == 0 ? w1 : nmid == 1 ? w2 : nmid == 2 ? w3 : w4;
    return;
  }
  if (order == 2) { /* getgpts.f:357-414 */
    const double g1 = 0.774596669241483, w1 = 0.555555555555556,
                 w2 = 3.555555555555556;
    static const int s[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, -1, -1},
                                {1, 1, -1},   {-1, -1, 1}, {-1, 1, 1},
                                {1, -1, 1},   {1, 1, 1}};
    if (gpn == 9) {
      *xi = *eta = *zeta = 0.0;
      *w = w2;
      return;
    }
    *xi = s[gpn - 1][0] * g1;
    *eta = s[gpn - 1][1] * g1;
    *zeta = s[gpn - 1][2] * g1;
    *w = w1;
    return;
  }
  if (order == 8) { /* getgpts.f:417-470, note the shorter constant */
    const double l1 = 0.577350269189626;
    static const int s[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, -1, -1},
                                {1, 1, -1},   {-1, -1, 1}, {-1, 1, 1},
                                {1, -1, 1},   {1, 1, 1}};
    *xi = s[gpn - 1][0] * l1;
    *eta = s[gpn - 1][1] * l1;
    *zeta = s[gpn - 1][2] * l1;
    *w = 1.0;
    return;
  }
  if (order == 9) { /* 14-point rule, getgpts.f:472-560 */
    const double l1 = 0.795822426, l2 = 0.758786911, l3 = 0.0,
                 w1 = 0.886426593, w2 = 0.335180055;
    static const double p[14][3] = {
        {-1, 0, 0}, {1, 0, 0},  {0, -1, 0},  {0, 1, 0},  {0, 0, -1},
        {0, 0, 1},  {-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1},
        {1, -1, 1}, {1, -1, -1}, {1, 1, -1},  {1, 1, 1}};
    int k = gpn - 1;
    double l = k < 6 ? l1 : l2;
    *xi = p[k][0] == 0 ? l3 : p[k][0] * l;
    *eta = p[k][1] == 0 ? l3 : p[k][1] * l;
    *zeta = p[k][2] == 0 ? l3 : p[k][2] * l;
    *w = k < 6 ? w1 : w2;
    return;
  }
}

/* quad2, getgpts.f:642 -- 8-node brick 2x2x2 (eta fastest, then xi). */
static void quad2(int gpn, double* xi, double* eta, double* zeta, double* w) {
  const double l1 = 0.5773502691896257645;
  static const int s[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, -1, -1},
                              {1, 1, -1},   {-1, -1, 1}, {-1, 1, 1},
                              {1, -1, 1},   {1, 1, 1}};
  *xi = s[gpn - 1][0] * l1;
  *eta = s[gpn - 1][1] * l1;
  *zeta = s[gpn - 1][2] * l1;
  *w = 1.0;
}

/* quad6, getgpts.f:771 -- tetrahedron rules in volume coordinates; returns
   (s2, s3, s4); the weights sum to 1 and the 1/6 of the tetrahedron volume
   is applied by adjust_scalar_weights (initst.f:888). */
static void quad6(int order, int point, double* s2, double* s3, double* s4,
                  double* weight) {
  double s1 = 0, a, b, w;
  if (order == 1) {
    *s2 = *s3 = *s4 = 0.25;
    *weight = 1.0;
    return;
  }
  if (order == 4) {
    a = 0.58541020;
    b = 0.13819660;
    w = 0.25;
  } else { /* order 5 */
    if (point == 1) {
      a = 0.25;
      b = 0.25;
      w = -4.0 / 5.0;
    } else {
      a = 1.0 / 2.0;
      b = 1.0 / 6.0;
      w = 9.0 / 20.0;
    }
    if (point == 1) {
      *s2 = *s3 = *s4 = 0.25;
      *weight = w;
      return;
    }
    point -= 1;
  }
  (void)s1;
  *s2 = point == 2 ? a : b;
  *s3 = point == 3 ? a : b;
  *s4 = point == 4 ? a : b;
  *weight = w;
}

void orc_getgpts(int etype, int order, int gpn, double* xi, double* eta,
                 double* zeta, double* w) {
  switch (etype) {
    case W3D_Q3DISOP: quad1(order, gpn, xi, eta, zeta, w); break;
    case W3D_L3DISOP: quad2(gpn, xi, eta, zeta, w); break;
    case W3D_TET10: quad6(order, gpn, xi, eta, zeta, w); break;
    default: *xi = *eta = *zeta = *w = 0.0;
  }
}

/* deriv1, derivs.f:182 -- 20-node serendipity brick. */
static void deriv1(double r, double s, double t, double* nr, double* ns,
                   double* nt) {
  const double one = 1.0, half = 0.5, fourth = 0.25, eighth = 0.125;
  double rp1 = one + r, rm1 = one - r, sp1 = one + s, sm1 = one - s,
         tp1 = one + t, tm1 = one - t, rps = r + s, rms = r - s, spt = s + t,
         smt = s - t;
  nr[0] = sm1 * (rp1 + rps - t) * tp1 * eighth;
  nr[1] = sm1 * (rp1 + rps + t) * tm1 * eighth;
  nr[2] = sp1 * (rp1 + rms + t) * tm1 * eighth;
  nr[3] = sp1 * (rp1 + rms - t) * tp1 * eighth;
  nr[4] = -sm1 * (rm1 - rms - t) * tp1 * eighth;
  nr[5] = -sm1 * (rm1 - rms + t) * tm1 * eighth;
  nr[6] = -sp1 * (rm1 - rps + t) * tm1 * eighth;
  nr[7] = -sp1 * (rm1 - rps - t) * tp1 * eighth;
  nr[8] = -sm1 * tm1 * tp1 * fourth;
  nr[9] = -sm1 * sp1 * tm1 * fourth;
  nr[10] = -sp1 * tm1 * tp1 * fourth;
  nr[11] = -sm1 * sp1 * tp1 * fourth;
  nr[12] = sm1 * tm1 * tp1 * fourth;
  nr[13] = sm1 * sp1 * tm1 * fourth;
  nr[14] = sp1 * tm1 * tp1 * fourth;
  nr[15] = sm1 * sp1 * tp1 * fourth;
  nr[16] = -r * sm1 * tp1 * half;
  nr[17] = -r * sm1 * tm1 * half;
  nr[18] = -r * sp1 * tm1 * half;
  nr[19] = -r * sp1 * tp1 * half;

  ns[0] = rm1 * (rp1 + s + smt) * tp1 * eighth;
  ns[1] = rm1 * (rp1 + s + spt) * tm1 * eighth;
  ns[2] = -rm1 * (rp1 - s - smt) * tm1 * eighth;
  ns[3] = -rm1 * (rp1 - s - spt) * tp1 * eighth;
  ns[4] = rp1 * (rm1 + s + smt) * tp1 * eighth;
  ns[5] = rp1 * (rm1 + s + spt) * tm1 * eighth;
  ns[6] = -rp1 * (rm1 - s - smt) * tm1 * eighth;
  ns[7] = -rp1 * (rm1 - s - spt) * tp1 * eighth;
  ns[8] = -rm1 * tm1 * tp1 * fourth;
  ns[9] = -rm1 * s * tm1 * half;
  ns[10] = rm1 * tm1 * tp1 * fourth;
  ns[11] = -rm1 * s * tp1 * half;
  ns[12] = -rp1 * tm1 * tp1 * fourth;
  ns[13] = -rp1 * s * tm1 * half;
  ns[14] = rp1 * tm1 * tp1 * fourth;
  ns[15] = -rp1 * s * tp1 * half;
  ns[16] = -rm1 * rp1 * tp1 * fourth;
  ns[17] = -rm1 * rp1 * tm1 * fourth;
  ns[18] = rm1 * rp1 * tm1 * fourth;
  ns[19] = rm1 * rp1 * tp1 * fourth;

  nt[0] = -rm1 * sm1 * (rp1 + smt - t) * eighth;
  nt[1] = rm1 * sm1 * (rp1 + spt + t) * eighth;
  nt[2] = rm1 * sp1 * (rp1 - smt + t) * eighth;
  nt[3] = -rm1 * sp1 * (rp1 - spt - t) * eighth;
  nt[4] = -rp1 * sm1 * (rm1 + smt - t) * eighth;
  nt[5] = rp1 * sm1 * (rm1 + spt + t) * eighth;
  nt[6] = rp1 * sp1 * (rm1 - smt + t) * eighth;
  nt[7] = -rp1 * sp1 * (rm1 - spt - t) * eighth;
  nt[8] = -rm1 * sm1 * t * half;
  nt[9] = -rm1 * sm1 * sp1 * fourth;
  nt[10] = -rm1 * sp1 * t * half;
  nt[11] = rm1 * sm1 * sp1 * fourth;
  nt[12] = -rp1 * sm1 * t * half;
  nt[13] = -rp1 * sm1 * sp1 * fourth;
  nt[14] = -rp1 * sp1 * t * half;
  nt[15] = rp1 * sm1 * sp1 * fourth;
  nt[16] = rm1 * rp1 * sm1 * fourth;
  nt[17] = -rm1 * rp1 * sm1 * fourth;
  nt[18] = -rm1 * rp1 * sp1 * fourth;
  nt[19] = rm1 * rp1 * sp1 * fourth;
}

/* deriv2, derivs.f:283 -- 8-node brick.  Node k sits at
   (xi,eta,zeta) = (-,-,+) (-,-,-) (-,+,-) (-,+,+) (+,-,+) (+,-,-) (+,+,-)
   (+,+,+) for k = 1..8. */
static void deriv2(double xi, double eta, double zeta, double* nxi,
                   double* neta, double* nzeta) {
  const double one = 1.0, one25 = 0.125;
  double xp = one + xi, ep = one + eta, zp = one + zeta, xm = one - xi,
         em = one - eta, zm = one - zeta;
  nxi[0] = -em * zp * one25;
  neta[0] = -xm * zp * one25;
  nzeta[0] = xm * em * one25;
  nxi[1] = -em * zm * one25;
  neta[1] = -xm * zm * one25;
  nzeta[1] = -xm * em * one25;
  nxi[2] = -ep * zm * one25;
  neta[2] = xm * zm * one25;
  nzeta[2] = -xm * ep * one25;
  nxi[3] = -ep * zp * one25;
  neta[3] = xm * zp * one25;
  nzeta[3] = xm * ep * one25;
  nxi[4] = em * zp * one25;
  neta[4] = -xp * zp * one25;
  nzeta[4] = xp * em * one25;
  nxi[5] = em * zm * one25;
  neta[5] = -xp * zm * one25;
  nzeta[5] = -xp * em * one25;
  nxi[6] = ep * zm * one25;
  neta[6] = xp * zm * one25;
  nzeta[6] = -xp * ep * one25;
  nxi[7] = ep * zp * one25;
  neta[7] = xp * zp * one25;
  nzeta[7] = xp * ep * one25;
}

/* deriv6, derivs.f:606 -- 10-node tetrahedron in (s2,s3,s4). */
static void deriv6(double s2, double s3, double s4, double* q2, double* q3,
                   double* q4) {
  const double zero = 0.0, one = 1.0, four = 4.0;
  double s1 = one - s2 - s3 - s4;
  q2[0] = one - four * s1;
  q2[1] = four * s2 - one;
  q2[2] = zero;
  q2[3] = zero;
  q2[4] = four * (s1 - s2);
  q2[5] = four * s3;
  q2[6] = -four * s3;
  q2[7] = -four * s4;
  q2[8] = four * s4;
  q2[9] = zero;
  q3[0] = one - four * s1;
  q3[1] = zero;
  q3[2] = four * s3 - one;
  q3[3] = zero;
  q3[4] = -four * s2;
  q3[5] = four * s2;
  q3[6] = four * (s1 - s3);
  q3[7] = -four * s4;
  q3[8] = zero;
  q3[9] = four * s4;
  q4[0] = one - four * s1;
  q4[1] = zero;
  q4[2] = zero;
  q4[3] = four * s4 - one;
  q4[4] = -four * s2;
  q4[5] = zero;
  q4[6] = -four * s3;
  q4[7] = four * (s1 - s4);
  q4[8] = four * s2;
  q4[9] = four * s3;
}

void orc_derivs(int etype, double xi, double eta, double zeta, double* nxi,
                double* neta, double* nzeta) {
  switch (etype) {
    case W3D_Q3DISOP: deriv1(xi, eta, zeta, nxi, neta, nzeta); break;
    case W3D_L3DISOP: deriv2(xi, eta, zeta, nxi, neta, nzeta); break;
    case W3D_TET10: deriv6(xi, eta, zeta, nxi, neta, nzeta); break;
    default: break;
  }
}

/* Shape-function values (shapef.f:15): shape1 (:169, 20-node brick; the
   corner functions as -(1-+r)(1-+s)(...)(1-+t)/8 with the factors in that
   order), shape2 (:223, 8-node brick), shape6 (:447, 10-node tetrahedron
   in (s2, s3, s4), s1 = 1 - s2 - s3 - s4).  Node numbering as derivs. */
void orc_shapef(int etype, double xi, double eta, double zeta, double* n) {
  const double one = 1.0, two = 2.0, four = 4.0, fourth = 0.25,
               eighth = 0.125;
  if (etype == W3D_L3DISOP) {
    double xp = one + xi, ep = one + eta, zp = one + zeta;
    double xm = one - xi, em = one - eta, zm = one - zeta;
    n[0] = xm * em * zp * eighth;
    n[1] = xm * em * zm * eighth;
    n[2] = xm * ep * zm * eighth;
    n[3] = xm * ep * zp * eighth;
    n[4] = xp * em * zp * eighth;
    n[5] = xp * em * zm * eighth;
    n[6] = xp * ep * zm * eighth;
    n[7] = xp * ep * zp * eighth;
  } else if (etype == W3D_Q3DISOP) {
    double rp = one + xi, rm = one - xi, sp = one + eta, sm = one - eta,
           tp = one + zeta, tm = one - zeta, t = zeta;
    /* corners: sign pattern of (r, s, t) per node, the mid-side nodes
       per the edge they sit on */
    n[0] = -rm * sm * (rp + sp - t) * tp * eighth;
    n[1] = -rm * sm * (rp + sp + t) * tm * eighth;
    n[2] = -rm * sp * (rp + sm + t) * tm * eighth;
    n[3] = -rm * sp * (rp + sm - t) * tp * eighth;
    n[4] = -rp * sm * (rm + sp - t) * tp * eighth;
    n[5] = -rp * sm * (rm + sp + t) * tm * eighth;
    n[6] = -rp * sp * (rm + sm + t) * tm * eighth;
    n[7] = -rp * sp * (rm + sm - t) * tp * eighth;
    n[8] = rm * sm * tm * tp * fourth;
    n[9] = rm * sm * sp * tm * fourth;
    n[10] = rm * sp * tm * tp * fourth;
    n[11] = rm * sm * sp * tp * fourth;
    n[12] = rp * sm * tm * tp * fourth;
    n[13] = rp * sm * sp * tm * fourth;
    n[14] = rp * sp * tm * tp * fourth;
    n[15] = rp * sm * sp * tp * fourth;
    n[16] = rm * rp * sm * tp * fourth;
    n[17] = rm * rp * sm * tm * fourth;
    n[18] = rm * rp * sp * tm * fourth;
    n[19] = rm * rp * sp * tp * fourth;
  } else if (etype == W3D_TET10) {
    double s2 = xi, s3 = eta, s4 = zeta, s1 = one - s2 - s3 - s4;
    n[0] = two * s1 * s1 - s1;
    n[1] = two * s2 * s2 - s2;
    n[2] = two * s3 * s3 - s3;
    n[3] = two * s4 * s4 - s4;
    n[4] = four * s1 * s2;
    n[5] = four * s2 * s3;
    n[6] = four * s1 * s3;
    n[7] = four * s1 * s4;
    n[8] = four * s2 * s4;
    n[9] = four * s3 * s4;
  }
}

/* jacob1, gtmat1.f:300 (3-D solid branch): J(r,c) = sum_j dN_j/dxi_r x_j,c;
   determinant by cofactors of row 1; gama = J^-1 by the adjugate.  jac and
   gama are row-major [r][c]. */
void orc_jacob1(int nnode, const double* nxi, const double* neta,
                const double* nzeta, const double* ce, double jac[9],
                double* dj, double gama[9]) {
  for (int k = 0; k < 9; ++k) jac[k] = 0.0;
  for (int j = 0; j < nnode; ++j) {
    double x = ce[j], y = ce[nnode + j], z = ce[2 * nnode + j];
    jac[0] = jac[0] + nxi[j] * x;
    jac[1] = jac[1] + nxi[j] * y;
    jac[2] = jac[2] + nxi[j] * z;
    jac[3] = jac[3] + neta[j] * x;
    jac[4] = jac[4] + neta[j] * y;
    jac[5] = jac[5] + neta[j] * z;
    jac[6] = jac[6] + nzeta[j] * x;
    jac[7] = jac[7] + nzeta[j] * y;
    jac[8] = jac[8] + nzeta[j] * z;
  }
#define J(r, c) jac[3 * ((r)-1) + (c)-1]
  double j1 = J(2, 2) * J(3, 3) - J(2, 3) * J(3, 2);
  double j2 = J(2, 1) * J(3, 3) - J(2, 3) * J(3, 1);
  double j3 = J(2, 1) * J(3, 2) - J(2, 2) * J(3, 1);
  double d = J(1, 1) * j1 - J(1, 2) * j2 + J(1, 3) * j3;
  *dj = d;
  gama[0] = j1 / d;
  gama[3] = -j2 / d;
  gama[6] = j3 / d;
  gama[1] = (J(3, 2) * J(1, 3) - J(1, 2) * J(3, 3)) / d;
  gama[4] = (J(1, 1) * J(3, 3) - J(3, 1) * J(1, 3)) / d;
  gama[7] = (J(1, 2) * J(3, 1) - J(1, 1) * J(3, 2)) / d;
  gama[2] = (J(1, 2) * J(2, 3) - J(1, 3) * J(2, 2)) / d;
  gama[5] = (J(1, 3) * J(2, 1) - J(1, 1) * J(2, 3)) / d;
  gama[8] = (J(1, 1) * J(2, 2) - J(1, 2) * J(2, 1)) / d;
#undef J
}

/*
 * blcmp1, blcmp1.f:15: the linear B (rows xx, yy, zz, xy, yz, xz; dof
 * columns u-block 1..N, v-block N+1..2N, w-block 2N+1..3N), and bmod,
 * bmod.f:16, the mean-dilatation B-bar of the 8-node brick when vol is
 * given (vol is the (8,3) volume-averaged dN/dx of vol_terms/vol_avg,
 * node-major).  b is row-major [r][dof], fully overwritten.
 */
void orc_bmatrix(int nnode, const double gama[9], const double* nxi,
                 const double* neta, const double* nzeta, const double* vol,
                 double eps_bbar, double* b) {
  int T = 3 * nnode;
  memset(b, 0, sizeof(double) * 6 * T);
  for (int j = 0; j < nnode; ++j) {
    double bx = gama[0] * nxi[j] + gama[1] * neta[j] + gama[2] * nzeta[j];
    double by = gama[3] * nxi[j] + gama[4] * neta[j] + gama[5] * nzeta[j];
    double bz = gama[6] * nxi[j] + gama[7] * neta[j] + gama[8] * nzeta[j];
    b[0 * T + j] = bx;
    b[3 * T + j] = by;
    b[5 * T + j] = bz;
    b[1 * T + nnode + j] = by;
    b[3 * T + nnode + j] = bx;
    b[4 * T + nnode + j] = bz;
    b[2 * T + 2 * nnode + j] = bz;
    b[4 * T + 2 * nnode + j] = by;
    b[5 * T + 2 * nnode + j] = bx;
  }
  if (!vol) return;
  /* bmod: alpha = (2 + eps)/3, beta = (1 - eps)/3 */
  const double one = 1.0, two = 2.0, third = 1.0 / 3.0;
  double alpha = (two + eps_bbar) * third;
  double beta = (one - eps_bbar) * third;
  double* r1 = b;
  double* r2 = b + T;
  double* r3 = b + 2 * T;
  for (int j = 0; j < 8; ++j) {
    double d = r1[j];
    r1[j] = alpha * d + beta * vol[3 * j + 0];
    r2[j] = beta * (vol[3 * j + 0] - d);
  }
  for (int j = 0; j < 8; ++j) {
    r1[8 + j] = beta * (vol[3 * j + 1] - r2[8 + j]);
    r1[16 + j] = beta * (vol[3 * j + 2] - r3[16 + j]);
    r2[8 + j] = alpha * r2[8 + j] + beta * vol[3 * j + 1];
    r2[16 + j] = r1[16 + j];
    r3[j] = r2[j];
    r3[8 + j] = r1[8 + j];
    r3[16 + j] = alpha * r3[16 + j] + beta * vol[3 * j + 2];
  }
}
