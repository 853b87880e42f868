// w3d_elements.cpp -- host-side Gauss rules and parametric shape-function
// derivatives; they fill the GaussTable that the kernels stage into LDS.
//
// Rules and node numbering (file:line under the reference's src/):
//   hex20 (q3disop) quad1 getgpts.f:177 / deriv1 derivs.f:182
//   hex8  (l3disop) quad2 getgpts.f:642 / deriv2 derivs.f:283
//   tet10           quad6 getgpts.f:771 / deriv6 derivs.f:606
// The constants are the reference's, including its truncated abscissae.
#include <cmath>
#include <cstring>

#include "w3d_internal.h"

namespace w3d {

int ngp_of(int etype, int order) {
  if (etype == W3D_L3DISOP) return order == 1 ? 8 : 0;
  if (etype == W3D_Q3DISOP)
    return order == 1 ? 27 : order == 2 ? 9 : order == 8 ? 8 : order == 9 ? 14
                                                                          : 0;
  if (etype == W3D_TET10)
    return (order == 1 || order == 4 || order == 5) ? order : 0;
  return 0;
}

int nnode_of(int etype) {
  return etype == W3D_Q3DISOP ? 20 : etype == W3D_L3DISOP ? 8
                                   : etype == W3D_TET10   ? 10
                                                          : 0;
}

static const int kCube[8][3] = {{-1, -1, -1}, {-1, 1, -1}, {1, -1, -1},
                                {1, 1, -1},   {-1, -1, 1}, {-1, 1, 1},
                                {1, -1, 1},   {1, 1, 1}};

static void point(int etype, int order, int g, double p[3], double* w) {
  if (etype == W3D_L3DISOP) {
    const double l = 0.5773502691896257645;
    for (int k = 0; k < 3; ++k) p[k] = kCube[g][k] * l;
    *w = 1.0;
    return;
  }
  if (etype == W3D_Q3DISOP) {
    if (order == 1) {  // 3x3x3: eta fastest, then xi, then zeta
      const double l1 = 0.774596669241483;
      const double lv[3] = {-l1, 0.0, l1};
      const double wv[4] = {0.1714677640603571, 0.2743484224965712,
                            0.4389574759945135, 0.7023319615912210};
      int ie = g % 3, ix = (g / 3) % 3, iz = g / 9;
      p[0] = lv[ix];
      p[1] = lv[ie];
      p[2] = lv[iz];
      *w = wv[(ix == 1) + (ie == 1) + (iz == 1)];
    } else if (order == 8) {
      const double l = 0.577350269189626;
      for (int k = 0; k < 3; ++k) p[k] = kCube[g][k] * l;
      *w = 1.0;
    } else if (order == 2) {
      const double l = 0.774596669241483;
      if (g == 8) {
        p[0] = p[1] = p[2] = 0.0;
        *w = 3.555555555555556;
      } else {
        for (int k = 0; k < 3; ++k) p[k] = kCube[g][k] * l;
        *w = 0.555555555555556;
      }
    } else {  // order 9, 14 points
      static const int s[14][3] = {
          {-1, 0, 0},  {1, 0, 0},   {0, -1, 0}, {0, 1, 0},  {0, 0, -1},
          {0, 0, 1},   {-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1},
          {1, -1, 1},  {1, -1, -1}, {1, 1, -1}, {1, 1, 1}};
      const double l = g < 6 ? 0.795822426 : 0.758786911;
      for (int k = 0; k < 3; ++k) p[k] = s[g][k] == 0 ? 0.0 : s[g][k] * l;
      *w = g < 6 ? 0.886426593 : 0.335180055;
    }
    return;
  }
  // tet10: volume coordinates (s2, s3, s4)
  if (order == 1) {
    p[0] = p[1] = p[2] = 0.25;
    *w = 1.0;
    return;
  }
  double a, b;
  int pt = g + 1;
  if (order == 4) {
    a = 0.58541020;
    b = 0.13819660;
    *w = 0.25;
  } else {
    if (pt == 1) {
      p[0] = p[1] = p[2] = 0.25;
      *w = -4.0 / 5.0;
      return;
    }
    a = 1.0 / 2.0;
    b = 1.0 / 6.0;
    *w = 9.0 / 20.0;
    pt -= 1;
  }
  p[0] = pt == 2 ? a : b;
  p[1] = pt == 3 ? a : b;
  p[2] = pt == 4 ? a : b;
}

static void derivs(int etype, const double p[3], double* nx, double* ny,
                   double* nz) {
  if (etype == W3D_L3DISOP) {
    // 8-node brick: node natural coordinates (derivs.f:283 numbering)
    static const int nat[8][3] = {{-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1},
                                  {-1, 1, 1},  {1, -1, 1},   {1, -1, -1},
                                  {1, 1, -1},  {1, 1, 1}};
    for (int j = 0; j < 8; ++j) {
      double a = 1.0 + nat[j][0] * p[0], b = 1.0 + nat[j][1] * p[1],
             c = 1.0 + nat[j][2] * p[2];
      nx[j] = nat[j][0] * b * c * 0.125;
      ny[j] = nat[j][1] * a * c * 0.125;
      nz[j] = nat[j][2] * a * b * 0.125;
    }
    return;
  }
  if (etype == W3D_Q3DISOP) {
    // 20-node serendipity brick, derivs.f:182 numbering
    static const int nat[20][3] = {
        {-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, 1},
        {1, -1, -1}, {1, 1, -1},   {1, 1, 1},   {-1, -1, 0}, {-1, 0, -1},
        {-1, 1, 0},  {-1, 0, 1},   {1, -1, 0},  {1, 0, -1}, {1, 1, 0},
        {1, 0, 1},   {0, -1, 1},   {0, -1, -1}, {0, 1, -1}, {0, 1, 1}};
    const double r = p[0], s = p[1], t = p[2];
    for (int j = 0; j < 20; ++j) {
      int ri = nat[j][0], si = nat[j][1], ti = nat[j][2];
      if (ri != 0 && si != 0 && ti != 0) {
        double a = 1.0 + ri * r, b = 1.0 + si * s, c = 1.0 + ti * t;
        double q = ri * r + si * s + ti * t - 2.0;
        nx[j] = 0.125 * ri * b * c * (q + a);
        ny[j] = 0.125 * si * a * c * (q + b);
        nz[j] = 0.125 * ti * a * b * (q + c);
      } else if (ri == 0) {
        double b = 1.0 + si * s, c = 1.0 + ti * t;
        nx[j] = -0.5 * r * b * c;
        ny[j] = 0.25 * si * (1.0 - r * r) * c;
        nz[j] = 0.25 * ti * (1.0 - r * r) * b;
      } else if (si == 0) {
        double a = 1.0 + ri * r, c = 1.0 + ti * t;
        nx[j] = 0.25 * ri * (1.0 - s * s) * c;
        ny[j] = -0.5 * s * a * c;
        nz[j] = 0.25 * ti * (1.0 - s * s) * a;
      } else {
        double a = 1.0 + ri * r, b = 1.0 + si * s;
        nx[j] = 0.25 * ri * (1.0 - t * t) * b;
        ny[j] = 0.25 * si * (1.0 - t * t) * a;
        nz[j] = -0.5 * t * a * b;
      }
    }
    return;
  }
  // tet10 in (s2, s3, s4) with s1 = 1 - s2 - s3 - s4
  const double s2 = p[0], s3 = p[1], s4 = p[2], s1 = 1.0 - s2 - s3 - s4;
  const double d2[10] = {1 - 4 * s1, 4 * s2 - 1, 0, 0, 4 * (s1 - s2),
                         4 * s3,     -4 * s3,    -4 * s4, 4 * s4, 0};
  const double d3[10] = {1 - 4 * s1, 0, 4 * s3 - 1, 0, -4 * s2,
                         4 * s2,     4 * (s1 - s3), -4 * s4, 0, 4 * s4};
  const double d4[10] = {1 - 4 * s1, 0, 0, 4 * s4 - 1, -4 * s2,
                         0,          -4 * s3, 4 * (s1 - s4), 4 * s2, 4 * s3};
  std::memcpy(nx, d2, sizeof d2);
  std::memcpy(ny, d3, sizeof d3);
  std::memcpy(nz, d4, sizeof d4);
}

// Shape-function values at p (shapef.f:15: shape2 :223, shape1 :169,
// shape6 :447), node numbering as derivs.
static void shapes(int etype, const double p[3], double* sf) {
  if (etype == W3D_L3DISOP) {
    static const int nat[8][3] = {{-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1},
                                  {-1, 1, 1},  {1, -1, 1},   {1, -1, -1},
                                  {1, 1, -1},  {1, 1, 1}};
    for (int j = 0; j < 8; ++j)
      sf[j] = (1.0 + nat[j][0] * p[0]) * (1.0 + nat[j][1] * p[1]) *
              (1.0 + nat[j][2] * p[2]) * 0.125;
    return;
  }
  if (etype == W3D_Q3DISOP) {
    static const int nat[20][3] = {
        {-1, -1, 1}, {-1, -1, -1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, 1},
        {1, -1, -1}, {1, 1, -1},   {1, 1, 1},   {-1, -1, 0}, {-1, 0, -1},
        {-1, 1, 0},  {-1, 0, 1},   {1, -1, 0},  {1, 0, -1}, {1, 1, 0},
        {1, 0, 1},   {0, -1, 1},   {0, -1, -1}, {0, 1, -1}, {0, 1, 1}};
    const double r = p[0], s = p[1], t = p[2];
    for (int j = 0; j < 20; ++j) {
      int ri = nat[j][0], si = nat[j][1], ti = nat[j][2];
      double a = 1.0 + ri * r, b = 1.0 + si * s, c = 1.0 + ti * t;
      if (ri != 0 && si != 0 && ti != 0)
        sf[j] = 0.125 * a * b * c * (ri * r + si * s + ti * t - 2.0);
      else if (ri == 0)
        sf[j] = 0.25 * (1.0 - r * r) * b * c;
      else if (si == 0)
        sf[j] = 0.25 * (1.0 - s * s) * a * c;
      else
        sf[j] = 0.25 * (1.0 - t * t) * a * b;
    }
    return;
  }
  const double s2 = p[0], s3 = p[1], s4 = p[2], s1 = 1.0 - s2 - s3 - s4;
  const double v[10] = {s1 * (2 * s1 - 1), s2 * (2 * s2 - 1),
                        s3 * (2 * s3 - 1), s4 * (2 * s4 - 1),
                        4 * s1 * s2,       4 * s2 * s3,
                        4 * s1 * s3,       4 * s1 * s4,
                        4 * s2 * s4,       4 * s3 * s4};
  std::memcpy(sf, v, sizeof v);
}

// Least-squares solution x (4) of min |A x - b|, A (m x 4) column-major,
// by Householder QR and back substitution (the steps of LAPACK's DGELS for
// a full-rank A); for m < 4 the minimum-norm solution A^T (A A^T)^-1 b of
// the one-point rule (m = 1).
static void lsq4(int m, const double* a_in, const double* b_in, double* x) {
  double a[4 * kMaxGp], b[kMaxGp];
  std::memcpy(a, a_in, sizeof(double) * 4 * m);
  std::memcpy(b, b_in, sizeof(double) * m);
  if (m < 4) {
    double s = 0.0;
    for (int j = 0; j < 4; ++j) s += a[j * m] * a[j * m];
    for (int j = 0; j < 4; ++j) x[j] = s > 0.0 ? a[j * m] * b[0] / s : 0.0;
    return;
  }
  for (int j = 0; j < 4; ++j) {
    double* v = a + j * m;
    double alpha = v[j], xn = 0.0;
    for (int i = j + 1; i < m; ++i) xn += v[i] * v[i];
    xn = std::sqrt(xn);
    double tau = 0.0;
    if (xn != 0.0) {
      const double beta = -std::copysign(std::sqrt(alpha * alpha + xn * xn),
                                         alpha);
      tau = (beta - alpha) / beta;
      const double sc = 1.0 / (alpha - beta);
      for (int i = j + 1; i < m; ++i) v[i] *= sc;
      v[j] = beta;
    }
    for (int k = j + 1; k <= 4; ++k) {
      double* y = k < 4 ? a + k * m : b;
      double w = y[j];
      for (int i = j + 1; i < m; ++i) w += v[i] * y[i];
      w *= tau;
      y[j] -= w;
      for (int i = j + 1; i < m; ++i) y[i] -= w * v[i];
    }
  }
  for (int j = 3; j >= 0; --j) {
    double s = b[j];
    for (int k = j + 1; k < 4; ++k) s -= a[k * m + j] * x[k];
    x[j] = s / a[j * m + j];
  }
}

bool fill_gauss_table(int etype, int order, GaussTable* t) {
  std::memset(t, 0, sizeof(*t));
  int g = ngp_of(etype, order), n = nnode_of(etype);
  if (g == 0 || n == 0) return false;
  t->etype = etype;
  t->order = order;
  t->ngp = g;
  t->nnode = n;
  double amat[4 * kMaxGp];
  for (int k = 0; k < g; ++k) {
    double p[3];
    point(etype, order, k, p, &t->w[k]);
    derivs(etype, p, t->dn[0][k], t->dn[1][k], t->dn[2][k]);
    shapes(etype, p, t->sf[k]);
    amat[k] = p[0];
    amat[g + k] = p[1];
    amat[2 * g + k] = p[2];
    amat[3 * g + k] = 1.0;
  }
  // columns of the fit's pseudo-inverse: the solution for unit data
  for (int k = 0; k < g; ++k) {
    double e[kMaxGp] = {0}, x[4];
    e[k] = 1.0;
    lsq4(g, amat, e, x);
    for (int d = 0; d < 3; ++d) t->lsq[d][k] = x[d];
  }
  return true;
}

}  // namespace w3d
