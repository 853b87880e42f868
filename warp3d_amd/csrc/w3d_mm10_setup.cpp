// w3d_mm10_setup.cpp -- host set-up of the mm10 crystal records: the
// crystal's elastic stiffness, its 12 slip systems and their rotation into
// the reference frame by the element's Euler angles.
//
// Reference routines restated (file:line under the reference's src/):
//   mod_crystals.f:421-600 finalize_new_crystal (fcc {111}<110> and bcc
//   {110}<111> slip systems), :1795-1840 (isotropic / cubic flexibility,
//   inverted by mm10_invsym), rknstr.f:1568-1640 setup_mm10_rknstr
//   (rotation g from the angles, Schmid tensors ms / qs of gT b (x) gT n,
//   stiffness rotated as Rstiff C Rstiff^T with Rstiff = rt2rve(gT)),
//   mm10_a.f:1310 mm10_rotation_matrix (Kocks convention), :1431
//   mm10_rt2rve.
#include <cmath>
#include <cstring>

#include "w3d_internal.h"

namespace w3d {
namespace {

// slip directions in units of 1/sqrt(2) and plane normals in units of
// 1/sqrt(3) of the fcc crystal (mod_crystals.f:441-500); bcc exchanges the
// two tables (mod_crystals.f:502-600)
const int kB[12][3] = {{0, -1, 1}, {1, 0, -1}, {-1, 1, 0}, {0, 1, 1},
                       {1, 0, 1},  {1, -1, 0}, {0, -1, 1}, {-1, 0, -1},
                       {1, 1, 0},  {0, 1, 1},  {1, 0, -1}, {-1, -1, 0}};
const int kN[12][3] = {{1, 1, 1},  {1, 1, 1},  {1, 1, 1},  {-1, -1, 1},
                       {-1, -1, 1}, {-1, -1, 1}, {-1, 1, 1}, {-1, 1, 1},
                       {-1, 1, 1}, {1, -1, 1}, {1, -1, 1}, {1, -1, 1}};

// inverse by LU with partial pivoting (column-major 6x6)
bool invert6(double* a) {
  double lu[36], x[36];
  std::memcpy(lu, a, sizeof lu);
  std::memset(x, 0, sizeof x);
  for (int i = 0; i < 6; ++i) x[i + 6 * i] = 1.0;
  int piv[6];
  for (int j = 0; j < 6; ++j) {
    int p = j;
    double amax = std::fabs(lu[j + 6 * j]);
    for (int i = j + 1; i < 6; ++i)
      if (std::fabs(lu[i + 6 * j]) > amax) {
        amax = std::fabs(lu[i + 6 * j]);
        p = i;
      }
    piv[j] = p;
    if (lu[p + 6 * j] == 0.0) return false;
    if (p != j)
      for (int k = 0; k < 6; ++k) std::swap(lu[j + 6 * k], lu[p + 6 * k]);
    double r = 1.0 / lu[j + 6 * j];
    for (int i = j + 1; i < 6; ++i) lu[i + 6 * j] *= r;
    for (int k = j + 1; k < 6; ++k) {
      double t = -lu[j + 6 * k];
      if (t != 0.0)
        for (int i = j + 1; i < 6; ++i) lu[i + 6 * k] += lu[i + 6 * j] * t;
    }
  }
  for (int c = 0; c < 6; ++c) {
    double* b = x + 6 * c;
    for (int j = 0; j < 6; ++j)
      if (piv[j] != j) std::swap(b[j], b[piv[j]]);
    for (int k = 0; k < 6; ++k)
      if (b[k] != 0.0)
        for (int i = k + 1; i < 6; ++i) b[i] -= b[k] * lu[i + 6 * k];
    for (int k = 5; k >= 0; --k)
      if (b[k] != 0.0) {
        b[k] /= lu[k + 6 * k];
        for (int i = 0; i < k; ++i) b[i] -= b[k] * lu[i + 6 * k];
      }
  }
  std::memcpy(a, x, sizeof x);
  return true;
}

// mm10_rt2rve (tensor-shear Voigt rotation), both column-major
void rt2rve(const double* rt, double* rv) {
  auto T = [&](int i, int j) { return rt[(i - 1) + 3 * (j - 1)]; };
  auto V = [&](int i, int j) -> double& { return rv[(i - 1) + 6 * (j - 1)]; };
  V(1, 1) = T(1, 1) * T(1, 1);
  V(1, 2) = T(1, 2) * T(1, 2);
  V(1, 3) = T(1, 3) * T(1, 3);
  V(1, 4) = 2.0 * T(1, 1) * T(1, 2);
  V(1, 5) = 2.0 * T(1, 3) * T(1, 2);
  V(1, 6) = 2.0 * T(1, 1) * T(1, 3);
  V(2, 1) = T(2, 1) * T(2, 1);
  V(2, 2) = T(2, 2) * T(2, 2);
  V(2, 3) = T(2, 3) * T(2, 3);
  V(2, 4) = 2.0 * T(2, 1) * T(2, 2);
  V(2, 5) = 2.0 * T(2, 3) * T(2, 2);
  V(2, 6) = 2.0 * T(2, 1) * T(2, 3);
  V(3, 1) = T(3, 1) * T(3, 1);
  V(3, 2) = T(3, 2) * T(3, 2);
  V(3, 3) = T(3, 3) * T(3, 3);
  V(3, 4) = 2.0 * T(3, 1) * T(3, 2);
  V(3, 5) = 2.0 * T(3, 3) * T(3, 2);
  V(3, 6) = 2.0 * T(3, 1) * T(3, 3);
  V(4, 1) = T(1, 1) * T(2, 1);
  V(4, 2) = T(1, 2) * T(2, 2);
  V(4, 3) = T(1, 3) * T(2, 3);
  V(4, 4) = T(1, 1) * T(2, 2) + T(2, 1) * T(1, 2);
  V(4, 5) = T(1, 2) * T(2, 3) + T(1, 3) * T(2, 2);
  V(4, 6) = T(1, 1) * T(2, 3) + T(1, 3) * T(2, 1);
  V(5, 1) = T(2, 1) * T(3, 1);
  V(5, 2) = T(3, 2) * T(2, 2);
  V(5, 3) = T(2, 3) * T(3, 3);
  V(5, 4) = T(2, 1) * T(3, 2) + T(2, 2) * T(3, 1);
  V(5, 5) = T(2, 2) * T(3, 3) + T(3, 2) * T(2, 3);
  V(5, 6) = T(2, 1) * T(3, 3) + T(2, 3) * T(3, 1);
  V(6, 1) = T(1, 1) * T(3, 1);
  V(6, 2) = T(1, 2) * T(3, 2);
  V(6, 3) = T(1, 3) * T(3, 3);
  V(6, 4) = T(1, 1) * T(3, 2) + T(1, 2) * T(3, 1);
  V(6, 5) = T(1, 2) * T(3, 3) + T(1, 3) * T(3, 2);
  V(6, 6) = T(1, 1) * T(3, 3) + T(3, 1) * T(1, 3);
}

}  // namespace

bool mm10_crystal_setup(const w3d_material& m, const double* angles,
                        CrystalDev* cr) {
  if ((m.slip_type != 1 && m.slip_type != 2) ||
      (m.elastic_type != 1 && m.elastic_type != 2))
    return false;
  const double e = m.cp_e, v = m.cp_nu, u = m.cp_mu;
  double c0[36];
  std::memset(c0, 0, sizeof c0);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) c0[i + 6 * j] = i == j ? 1 / e : -v / e;
  const double sh = m.elastic_type == 1 ? 2 * (1 + v) / e : 1 / u;
  for (int i = 3; i < 6; ++i) c0[i + 6 * i] = sh;
  if (!invert6(c0)) return false;
  // Kocks rotation (crystal -> reference)
  const double pi = 3.141592653589793, one_eighty = 180.0;
  double a = angles[0], b = angles[1], c = angles[2];
  if (m.angle_type == 1) {
    a = a * pi / one_eighty;
    b = b * pi / one_eighty;
    c = c * pi / one_eighty;
  }
  const double psi = a, th = b, phi = c;
  double* g = cr->g;
  g[0 + 3 * 0] = -std::sin(psi) * std::sin(phi) -
                 std::cos(psi) * std::cos(phi) * std::cos(th);
  g[0 + 3 * 1] = std::cos(psi) * std::sin(phi) -
                 std::sin(psi) * std::cos(phi) * std::cos(th);
  g[0 + 3 * 2] = std::cos(phi) * std::sin(th);
  g[1 + 3 * 0] = std::sin(psi) * std::cos(phi) -
                 std::cos(psi) * std::sin(phi) * std::cos(th);
  g[1 + 3 * 1] = -std::cos(psi) * std::cos(phi) -
                 std::sin(psi) * std::sin(phi) * std::cos(th);
  g[1 + 3 * 2] = std::sin(phi) * std::sin(th);
  g[2 + 3 * 0] = std::cos(psi) * std::sin(th);
  g[2 + 3 * 1] = std::sin(psi) * std::sin(th);
  g[2 + 3 * 2] = std::cos(th);
  double tr[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) tr[i + 3 * j] = g[j + 3 * i];
  const double fb = 1 / std::sqrt(2.0), fn = 1 / std::sqrt(3.0);
  for (int s = 0; s < 12; ++s) {
    double bv[3], nv[3], bs[3], ns[3];
    for (int k = 0; k < 3; ++k) {
      double x = fb * kB[s][k], y = fn * kN[s][k];
      if (m.slip_type == 1) {
        bv[k] = x;
        nv[k] = y;
      } else {
        bv[k] = y;
        nv[k] = x;
      }
    }
    for (int i = 0; i < 3; ++i) {
      bs[i] = tr[i] * bv[0] + tr[i + 3] * bv[1] + tr[i + 6] * bv[2];
      ns[i] = tr[i] * nv[0] + tr[i + 3] * nv[1] + tr[i + 6] * nv[2];
    }
    double A[9], As[9], Aa[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i + 3 * j] = bs[i] * ns[j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        As[i + 3 * j] = 0.5 * (A[i + 3 * j] + A[j + 3 * i]);
        Aa[i + 3 * j] = 0.5 * (A[i + 3 * j] - A[j + 3 * i]);
      }
    cr->ms[s][0] = As[0];
    cr->ms[s][1] = As[4];
    cr->ms[s][2] = As[8];
    cr->ms[s][3] = 2.0 * As[0 + 3 * 1];
    cr->ms[s][4] = 2.0 * As[1 + 3 * 2];
    cr->ms[s][5] = 2.0 * As[0 + 3 * 2];
    cr->qs[s][0] = Aa[1 + 3 * 2];
    cr->qs[s][1] = Aa[0 + 3 * 2];
    cr->qs[s][2] = Aa[0 + 3 * 1];
  }
  double rs[36], t66[36];
  rt2rve(tr, rs);
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) {
      double t = 0.0;
      for (int k = 0; k < 6; ++k) t = t + c0[i + 6 * k] * rs[j + 6 * k];
      t66[i + 6 * j] = t;
    }
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) {
      double t = 0.0;
      for (int k = 0; k < 6; ++k) t = t + rs[i + 6 * k] * t66[k + 6 * j];
      cr->C[i + 6 * j] = t;
    }
  for (int k = 0; k < 3; ++k) cr->angles[k] = angles[k];
  return true;
}

}  // namespace w3d
