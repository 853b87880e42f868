// w3d_mm10.hip -- gfx950 kernel of the mm10 crystal-plasticity update
// (Voce hardening, 12 fcc or bcc slip systems), one thread per Gauss point.
//
// Reference routines restated (file:line under the reference's src/):
//   rstgp1.f:4327 drive_10_update ([R n+1] of the polar decomposition for
//   finite strains, R = I for small strains; the iteration-0 "extrapolate
//   off" linear-elastic update), mm10_a.f:29 mm10 (crystal
//   loop, averages, store), :480 mm10_init_cc_hist0, :1777/:1808/:2605
//   step-1 history initialisation, :2493 mm10_copy_cc_hist, :1689
//   mm10_setup_np1, :852 mm10_setup, :2086 mm10_setup_voche, :1102
//   mm10_solve_crystal, :2657 mm10_solve_strup (sub-stepping), :2901
//   mm10_solve (predictor :3032, corrector :3224), :664 mm10_tangent,
//   :3991 symmetrisation, :3375 mm10_update_rotation, :3447 expmw3x3,
//   :3493 mm10_output, :1195 mm10_update_euler_angles, :998
//   mm10_store_cryhist; mm10_b.f:63 formR2, :177 formJ11, :281 formJ12,
//   :377 formJ21, :439 formJ22, :1073 formR1, :1169 form_dbarp, :1239
//   form_wbarp, :1312 form_wp, :1513 symsw, :1613 iw, :1813 slipinc,
//   :1845 rs, :1896-2074 the Voce functions (with ehard's extra
//   sign(slip) factor on its rate term, :1978-1983, kept as written).
//
// Every array of the local problem has compile-time indices, so the whole
// solve stays in registers; the pivoting of the small LU factorisations is
// done with selects instead of indexed row swaps.  Operation order follows
// the reference loop by loop (LAPACK/BLAS calls in their unblocked
// reference-BLAS order), except that each slip system's resolved shear is
// formed once per residual evaluation and reused (bitwise the value the
// reference recomputes).
#include <hip/hip_runtime.h>

#include "w3d_internal.h"

namespace w3d {
namespace {

constexpr int NS = 12;

struct Props {
  const CrystalDev* cr;
  double rate_n, tau_y, tau_v, voce_m, theta_0, iD_v, eps_dot_0_y;
  double atol, atol1, rtol, rtol1;
  int miter, alter_mode, angle_type;
};

// crystal_state of one crystal (the Voce subset).  The slip tensors
// (mm10_setup's ms, qc) depend only on the plastic rotation at n and R n+1,
// so every sub-step shares one copy.
// The rotated ms live in LDS, one column per lane (word w of lane t at
// g[w * kLanes + t]), so that the local problem's registers stay free; qs
// is formed where it is used from the crystal's reference vectors and the
// 3x3 rotation rw (mm10_rt2rvw), which keeps the LDS of a workgroup at
// 72 words a lane (four one-wave workgroups per compute unit).
constexpr int kLanes = 64;
constexpr int kGeoWords = 72;
struct CState {
  double D[6];
  double tt_rate, tinc, dg;
  const double* g;           // 72 ms words, lane-strided
  const double (*q0)[3];     // reference qs of the 12 systems
  double rw[9];              // RT2RVW(R Rp_n^T), column-major
  __device__ double ms(int i, int k) const { return g[(6 * i + k) * kLanes]; }
  // the lattice-spin vector of system i: rw q0[i] (mv3); qc of the
  // reference, or qs for a state whose rw is RT2RVW(Rp_n^T)
  __device__ void qs3(int i, double* q) const {
#pragma unroll
    for (int k = 0; k < 3; ++k)
      q[k] = rw[k] * q0[i][0] + rw[k + 3] * q0[i][1] + rw[k + 6] * q0[i][2];
  }
};

__device__ inline double sgn(double x) { return x >= 0.0 ? 1.0 : -1.0; }

// x^y of the Voce functions: by repeated squaring when y is a whole number
// in [0, 127] (the rate and Voce exponents of practice, e.g. n = 20, m = 1
// or 5: some ulps from the correctly rounded power, far inside the local
// solve's tolerances), ocml's pow otherwise.  y is a material constant, so
// the branch is uniform.
__device__ inline double vpow(double x, double y) {
  const int k = (int)y;
  if ((double)k == y && k >= 0 && k < 128) {
    double r = 1.0, b = x;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (k >> i & 1) r = r * b;
      if ((k >> i) > 1) b = b * b;
    }
    return r;
  }
  return pow(x, y);
}

// vpow over an array, in place: the exponent's bits are tested once for all
// N values (uniform branches around straight runs of N products); every
// value takes the products vpow takes.
template <int N>
__device__ inline void vpow_arr(double* v, double y) {
  const int k = (int)y;
  if ((double)k == y && k >= 0 && k < 128) {
    double r[N];
#pragma unroll
    for (int j = 0; j < N; ++j) r[j] = 1.0;
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      if (k >> i & 1) {
#pragma unroll
        for (int j = 0; j < N; ++j) r[j] = r[j] * v[j];
      }
      if ((k >> i) > 1) {
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = v[j] * v[j];
      }
    }
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = r[j];
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = pow(v[j], y);
  }
}

// [b] c with b 3x3 column-major
__device__ inline void mv3(double* a, const double* b, const double* c) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
    a[i] = b[i] * c[0] + b[i + 3] * c[1] + b[i + 6] * c[2];
}

// mm10_rt2rve of rt (3x3 column-major) -> rv (6x6 column-major)
__device__ inline void rt2rve(const double* rt, double* rv) {
#define RT(i, j) rt[((i)-1) + 3 * ((j)-1)]
#define RV(i, j) rv[((i)-1) + 6 * ((j)-1)]
  RV(1, 1) = RT(1, 1) * RT(1, 1);
  RV(1, 2) = RT(1, 2) * RT(1, 2);
  RV(1, 3) = RT(1, 3) * RT(1, 3);
  RV(1, 4) = 2.0 * RT(1, 1) * RT(1, 2);
  RV(1, 5) = 2.0 * RT(1, 3) * RT(1, 2);
  RV(1, 6) = 2.0 * RT(1, 1) * RT(1, 3);
  RV(2, 1) = RT(2, 1) * RT(2, 1);
  RV(2, 2) = RT(2, 2) * RT(2, 2);
  RV(2, 3) = RT(2, 3) * RT(2, 3);
  RV(2, 4) = 2.0 * RT(2, 1) * RT(2, 2);
  RV(2, 5) = 2.0 * RT(2, 3) * RT(2, 2);
  RV(2, 6) = 2.0 * RT(2, 1) * RT(2, 3);
  RV(3, 1) = RT(3, 1) * RT(3, 1);
  RV(3, 2) = RT(3, 2) * RT(3, 2);
  RV(3, 3) = RT(3, 3) * RT(3, 3);
  RV(3, 4) = 2.0 * RT(3, 1) * RT(3, 2);
  RV(3, 5) = 2.0 * RT(3, 3) * RT(3, 2);
  RV(3, 6) = 2.0 * RT(3, 1) * RT(3, 3);
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

// mm10_rt2rvw
__device__ inline void rt2rvw(const double* rt, double* rv) {
#define RW(i, j) rv[((i)-1) + 3 * ((j)-1)]
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

// LU with partial pivoting (first largest |pivot|) and the solve of
// nrhs right-hand sides, column-major with leading dimension N.  The
// eliminations and substitutions do not skip zero multipliers as LAPACK's
// loops do: across a wave the skip would be a divergent branch that saves
// nothing, and adding a zero product changes no value (signed zeros
// aside).
template <int N, int NRHS>
__device__ inline void lu_solve(double* a, double* b) {
  int piv[N];
  double rp[N];  // the pivots' reciprocals, for the back substitution
#pragma unroll
  for (int j = 0; j < N; ++j) {
    int p = j;
    double amax = fabs(a[j + N * j]);
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      double v = fabs(a[i + N * j]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    piv[j] = p;
#pragma unroll
    for (int i = j + 1; i < N; ++i) {
      if (i == p) {
#pragma unroll
        for (int k = 0; k < N; ++k) {
          double t = a[j + N * k];
          a[j + N * k] = a[i + N * k];
          a[i + N * k] = t;
        }
      }
    }
    const double r = 1.0 / a[j + N * j];
    rp[j] = r;
#pragma unroll
    for (int i = j + 1; i < N; ++i) a[i + N * j] = a[i + N * j] * r;
#pragma unroll
    for (int k = j + 1; k < N; ++k) {
      double t = -a[j + N * k];
#pragma unroll
      for (int i = j + 1; i < N; ++i)
        a[i + N * k] = a[i + N * k] + a[i + N * j] * t;
    }
  }
#pragma unroll
  for (int c = 0; c < NRHS; ++c) {
    double* x = b + N * c;
#pragma unroll
    for (int j = 0; j < N; ++j) {
#pragma unroll
      for (int i = j + 1; i < N; ++i) {
        if (i == piv[j]) {
          double t = x[j];
          x[j] = x[i];
          x[i] = t;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
#pragma unroll
      for (int i = k + 1; i < N; ++i) x[i] = x[i] - x[k] * a[i + N * k];
    }
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
      x[k] = x[k] * rp[k];
#pragma unroll
      for (int i = 0; i < k; ++i) x[i] = x[i] - x[k] * a[i + N * k];
    }
  }
}

// DPOSV('U', 6, 1) with the stiffness
__device__ inline void chol_solve6(const double* __restrict__ ain, double* b) {
  double a[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) a[k] = ain[k];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < j; ++k) s = s + a[k + 6 * j] * a[k + 6 * j];
    double ajj = sqrt(a[j + 6 * j] - s);
    a[j + 6 * j] = ajj;
#pragma unroll
    for (int c = j + 1; c < 6; ++c) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < j; ++k) t = t + a[k + 6 * j] * a[k + 6 * c];
      a[j + 6 * c] = (a[j + 6 * c] - t) / ajj;
    }
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    double t = b[j];
#pragma unroll
    for (int k = 0; k < j; ++k) t = t - a[k + 6 * j] * b[k];
    b[j] = t / a[j + 6 * j];
  }
#pragma unroll
  for (int j = 5; j >= 0; --j) {
    b[j] = b[j] / a[j + 6 * j];
#pragma unroll
    for (int i = 0; i < j; ++i) b[i] = b[i] - b[j] * a[i + 6 * j];
  }
}

// mm10_setup: effective strain increment (k_0 = 0: no geometric hardening)
__device__ inline void setup(const Props& p, CState& s) {
  const double twothirds = 2.0 / 3.0;
  double t1 = s.D[0] * s.D[0] + s.D[1] * s.D[1] + s.D[2] * s.D[2];
  double t2 = s.D[3] * s.D[3] + s.D[4] * s.D[4] + s.D[5] * s.D[5];
  s.dg = sqrt(twothirds * (t1 + 0.5 * t2));
  if (p.alter_mode) s.dg = p.eps_dot_0_y * s.tinc;
}

// mm10_a_mult_type_1: a = b c (3x3 column-major), the reference's order
__device__ inline void mul33(double* a, const double* b, const double* c) {
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      a[i + 3 * j] = b[i] * c[3 * j] + b[i + 3] * c[1 + 3 * j] +
                     b[i + 6] * c[2 + 3 * j];
}

// mm10_setup: the slip tensors rotated by the plastic rotation at n.  The
// rotation kept in the state (rw) is the one of qc, RT2RVW(R Rp_n^T)
// (mm10_a.f:891-892); with small strains (rot null, R = I) R Rp_n^T is
// Rp_n^T exactly and qc equals qs.  rot: the point's [R n+1], column-major
// at rot[k * SG].
__device__ inline void slip_geometry(const Props& p, const double* Rpn,
                                     double* gl, CState& s,
                                     const double* rot, int64_t SG) {
  double tn[9], RE[36];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) tn[i + 3 * j] = Rpn[j + 3 * i];
  rt2rve(tn, RE);
  if (rot) {
    double R[9], w[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = rot[k * SG];
    mul33(w, R, tn);
    rt2rvw(w, s.rw);
  } else {
    rt2rvw(tn, s.rw);
  }
  s.q0 = p.cr->qs;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double* m0 = p.cr->ms[i];
    double m[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) m[r] = RE[r] * m0[0];
#pragma unroll
    for (int j = 1; j < 6; ++j)
#pragma unroll
      for (int r = 0; r < 6; ++r) m[r] = m[r] + RE[r + 6 * j] * m0[j];
#pragma unroll
    for (int r = 0; r < 6; ++r) gl[(6 * i + r) * kLanes] = m[r];
  }
  s.g = gl;
}

__device__ inline double rs_of(const CState& s, const double* x, int i) {
  return x[0] * s.ms(i, 0) + x[1] * s.ms(i, 1) + x[2] * s.ms(i, 2) +
         x[3] * s.ms(i, 3) + x[4] * s.ms(i, 4) + x[5] * s.ms(i, 5);
}

__device__ inline void symsw(const double* s, const double* w, double* sw) {
  sw[0] = s[3] * w[2] - s[5] * w[1];
  sw[1] = s[3] * w[2] - s[4] * w[0];
  sw[2] = s[5] * w[1] + s[4] * w[0];
  sw[3] = 0.5 * (w[2] * (s[0] - s[1]) + w[0] * s[5] - w[1] * s[4]);
  sw[4] = 0.5 * (w[0] * (s[1] - s[2]) + w[1] * s[3] + w[2] * s[5]);
  sw[5] = 0.5 * (w[1] * (s[0] - s[2]) + w[0] * s[3] - w[2] * s[4]);
}

// resolved shears and slip increments at x (stress) and tau
__device__ inline void slips(const Props& p, const CState& s, const double* x,
                             double tt, double* rs, double* gam) {
  // one reciprocal of tau for the twelve quotients (a few ulps from the
  // quotients themselves, far inside the local solve's tolerances)
  const double itt = 1.0 / tt, dgt = s.dg * itt;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    rs[i] = rs_of(s, x, i);
    gam[i] = fabs(rs[i] * itt);
  }
  vpow_arr<NS>(gam, p.rate_n - 1.0);
#pragma unroll
  for (int i = 0; i < NS; ++i) gam[i] = dgt * gam[i] * rs[i];
}

// form_wp (with qc) and form_wbarp (with qs: a state whose rw is that
// of Rp_n^T)
__device__ inline void form_w(const Props& p, const CState& s,
                              const double* rs, const double* gam, double* w) {
  w[0] = w[1] = w[2] = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double q[3];
    s.qs3(i, q);
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = w[k] + gam[i] * q[k];
    double c = rs[i] * s.tinc * p.iD_v;
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = w[k] + c * q[k];
  }
}

// formR1 at x with tau tt, from the slips (rs, gam) at (x, tt)
__device__ inline void formR1_slips(const Props& p, const CState& s,
                                    const double* sn, const double* x,
                                    const double* rs, const double* gam,
                                    double* R1) {
  double dbar[6], wp[3], sw[6], w1[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) dbar[k] = 0.0;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double c = rs[i] * s.tinc * p.iD_v + gam[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) dbar[k] = dbar[k] + c * s.ms(i, k);
  }
  form_w(p, s, rs, gam, wp);
  symsw(x, wp, sw);
#pragma unroll
  for (int k = 0; k < 6; ++k) w1[k] = s.D[k] - dbar[k];
  const double* C = p.cr->C;
  double w2[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) w2[i] = C[i] * w1[0];
#pragma unroll
  for (int j = 1; j < 6; ++j)
#pragma unroll
    for (int i = 0; i < 6; ++i) w2[i] = w2[i] + C[i + 6 * j] * w1[j];
#pragma unroll
  for (int k = 0; k < 6; ++k) R1[k] = x[k] - sn[k] - w2[k] + 2.0 * sw[k];
}

// formR1 at x with tau tt
__device__ inline void formR1(const Props& p, const CState& s,
                              const double* sn, const double* x, double tt,
                              double* R1) {
  double rs[NS], gam[NS];
  slips(p, s, x, tt, rs, gam);
  formR1_slips(p, s, sn, x, rs, gam, R1);
}

// h_voche (tau_l = 0)
__device__ inline double h_voche(const Props& p, const double* gam, double tt,
                                 double tau_n) {
  double h = 0.0;
  double ht = 1.0 - (tt - p.tau_y) / p.tau_v + 0.0 / (tt - p.tau_y);
  double f = vpow(fabs(ht), p.voce_m) * sgn(ht);
#pragma unroll
  for (int i = 0; i < NS; ++i) h = h + f * fabs(gam[i]);
  return tau_n + p.theta_0 * h;
}

// formR (7 residuals); records the hardening rate
// (the reference evaluates the slips twice, in formR1 and here, from the
// same arguments: once here, bitwise the same values)
__device__ inline void formR(const Props& p, CState& s, const double* sn,
                             double tau_n, const double* x, double* R) {
  double rs[NS], gam[NS];
  slips(p, s, x, x[6], rs, gam);
  formR1_slips(p, s, sn, x, rs, gam, R);
  double h = h_voche(p, gam, x[6], tau_n);
  R[6] = x[6] - h;
  s.tt_rate = (h - tau_n) / s.tinc;
}

// formJ11 (6x6, column-major) from the slips (rs, gam) at (x, tt); with
// J12 also formJ's J12 column, from the same per-system vectors
// 2 (sym(x) w_i) + C ms_i that J11 takes
template <bool WITH_J12 = false>
__device__ inline void formJ11_slips(const Props& p, const CState& s,
                                     const double* x, double tt,
                                     const double* rs, const double* gam,
                                     double* J11, double* J12 = nullptr) {
  const double* C = p.cr->C;
  const double ptt = vpow(tt, p.rate_n);
#pragma unroll
  for (int k = 0; k < 36; ++k) J11[k] = 0.0;
  double dp[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) dp[i] = fabs(rs[i]);
  vpow_arr<NS>(dp, p.rate_n - 1.0);
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double d = dp[i];
    d = s.dg * p.rate_n / ptt * d;
    double dgdt = d + s.tinc * p.iD_v;
    double st[6], wv[6];
    double qv[3];
    s.qs3(i, qv);
    symsw(x, qv, st);
#pragma unroll
    for (int r = 0; r < 6; ++r) wv[r] = 2.0 * st[r];
#pragma unroll
    for (int j = 0; j < 6; ++j)
#pragma unroll
      for (int r = 0; r < 6; ++r) wv[r] = wv[r] + C[r + 6 * j] * s.ms(i, j);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double t = dgdt * s.ms(i, c);
#pragma unroll
      for (int r = 0; r < 6; ++r) J11[r + 6 * c] = J11[r + 6 * c] + wv[r] * t;
    }
    if (WITH_J12) {
      const double dgdh = -p.rate_n / tt * gam[i];
#pragma unroll
      for (int r = 0; r < 6; ++r) J12[r] = J12[r] + wv[r] * (1.0 * dgdh);
    }
  }
  double wp[3];
  form_w(p, s, rs, gam, wp);
  // + mm10_iw(wp), then + I
  J11[0 + 6 * 3] = J11[0 + 6 * 3] + 2.0 * wp[2];
  J11[0 + 6 * 5] = J11[0 + 6 * 5] + -2.0 * wp[1];
  J11[1 + 6 * 3] = J11[1 + 6 * 3] + 2.0 * wp[2];
  J11[1 + 6 * 4] = J11[1 + 6 * 4] + -2.0 * wp[0];
  J11[2 + 6 * 4] = J11[2 + 6 * 4] + 2.0 * wp[0];
  J11[2 + 6 * 5] = J11[2 + 6 * 5] + 2.0 * wp[1];
  J11[3 + 6 * 0] = J11[3 + 6 * 0] + wp[2];
  J11[3 + 6 * 1] = J11[3 + 6 * 1] + -wp[2];
  J11[3 + 6 * 4] = J11[3 + 6 * 4] + -wp[1];
  J11[3 + 6 * 5] = J11[3 + 6 * 5] + wp[0];
  J11[4 + 6 * 1] = J11[4 + 6 * 1] + wp[0];
  J11[4 + 6 * 2] = J11[4 + 6 * 2] + -wp[0];
  J11[4 + 6 * 3] = J11[4 + 6 * 3] + wp[1];
  J11[4 + 6 * 5] = J11[4 + 6 * 5] + wp[2];
  J11[5 + 6 * 0] = J11[5 + 6 * 0] + wp[1];
  J11[5 + 6 * 2] = J11[5 + 6 * 2] + -wp[1];
  J11[5 + 6 * 3] = J11[5 + 6 * 3] + wp[0];
  J11[5 + 6 * 4] = J11[5 + 6 * 4] + -wp[2];
  // the zero entries of iw are added too: x + 0.0 == x
#pragma unroll
  for (int i = 0; i < 6; ++i) J11[i + 6 * i] = J11[i + 6 * i] + 1.0;
}

__device__ inline void formJ11(const Props& p, const CState& s,
                               const double* x, double tt, double* J11) {
  double rs[NS], gam[NS];
  slips(p, s, x, tt, rs, gam);
  formJ11_slips(p, s, x, tt, rs, gam, J11);
}

// formJ: the 7x7 Jacobian (column-major, leading dimension 7)
__device__ inline void formJ(const Props& p, const CState& s, const double* x,
                             double* J) {
  const double tt = x[6];
  double rs[NS], gam[NS];
  slips(p, s, x, tt, rs, gam);  // once for J11 and the rest (bitwise)
  double J11[36];
  double J12[6] = {0, 0, 0, 0, 0, 0};
  formJ11_slips<true>(p, s, x, tt, rs, gam, J11, J12);
  const double ht = 1.0 - (tt - p.tau_y) / p.tau_v + 0.0 / (tt - p.tau_y);
  const double hm = vpow(fabs(ht), p.voce_m);
  double et[6] = {0, 0, 0, 0, 0, 0};
  double dp[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) dp[i] = fabs(rs[i]);
  vpow_arr<NS>(dp, p.rate_n - 2.0);
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    double c = hm * sgn(ht) * dp[i] * rs[i];
#pragma unroll
    for (int k = 0; k < 6; ++k) et[k] = et[k] + c * s.ms(i, k);
  }
  const double sc = p.theta_0 * s.dg * p.rate_n / vpow(tt, p.rate_n);
  double etau = 0.0;
  const double dty = tt - p.tau_y;
  const double iht = 1.0 / fabs(ht), ntt = p.rate_n / tt;
#pragma unroll
  for (int i = 0; i < NS; ++i)
    etau = etau + (p.voce_m * (-1.0 / p.tau_v - 0.0 / (dty * dty)) *
                       (fabs(gam[i]) * iht) -
                   fabs(gam[i]) * ntt * sgn(ht) * sgn(gam[i])) *
                      hm;
  etau = p.theta_0 * etau;
  etau = 1.0 - etau;
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int r = 0; r < 6; ++r) J[r + 7 * c] = J11[r + 6 * c];
#pragma unroll
  for (int r = 0; r < 6; ++r) J[r + 42] = J12[r];
#pragma unroll
  for (int c = 0; c < 6; ++c) J[6 + 7 * c] = -(sc * et[c]);
  J[48] = etau;
}

// (unrolled, so that the arrays it is given stay in registers)
template <int N>
__device__ inline bool any_nan(const double* x) {
  bool b = false;
#pragma unroll
  for (int i = 0; i < N; ++i) b |= isnan(x[i]);
  return b;
}

__device__ inline double dot6(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + a[3] * b[3] + a[4] * b[4] +
         a[5] * b[5];
}

// mm10_solve: predictor (6x6 Newton at fixed extrapolated tau) then the
// 7x7 corrector; Jg (49 words at stride jstride) receives the Jacobian of
// the last corrector iteration
__device__ __forceinline__ void solve(const Props& p, CState& s, const double* sn,
                      double tau_n, double tt_rate_n, const double* Dn,
                      int64_t dn_stride,
                      double* stress, double* tt, bool& fail, double dtinc,
                      double* Jg, int64_t jstride) {
  const double c = 1.0e-4, red = 0.5;
  const int mls = 10, mmin = 1;
  double x[7], inR1;
#pragma unroll
  for (int k = 0; k < 6; ++k) x[k] = stress[k];
  x[6] = *tt;
  {  // predictor
    double x1[6], R1[6], d1[6], d2[6], t1, t2, dtrace;
#pragma unroll
    for (int k = 0; k < 6; ++k) x1[k] = x[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) d1[k] = s.D[k];
    dtrace = (d1[0] + d1[1] + d1[2]) / 3.0;
    d1[0] = d1[0] - dtrace;
    d1[1] = d1[1] - dtrace;
    d1[2] = d1[2] - dtrace;
    t1 = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    t2 = d1[3] * d1[3] + d1[4] * d1[4] + d1[5] * d1[5];
    if (t1 + t2 == 0.0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) d1[k] = 0.0;
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) d1[k] = d1[k] / sqrt(t1 + t2);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) d2[k] = Dn ? Dn[k * dn_stride] : 0.0;
    dtrace = (d2[0] + d2[1] + d2[2]) / 3.0;
    d2[0] = d2[0] - dtrace;
    d2[1] = d2[1] - dtrace;
    d2[2] = d2[2] - dtrace;
    t1 = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    t2 = d2[3] * d2[3] + d2[4] * d2[4] + d2[5] * d2[5];
    if (t1 + t2 == 0.0) {
#pragma unroll
      for (int k = 0; k < 6; ++k) d2[k] = 0.0;
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) d2[k] = d2[k] / sqrt(t1 + t2);
    }
    t1 = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    t2 = d1[3] * d2[3] + d1[4] * d2[4] + d1[5] * d2[5];
    double cos_ang = t1 + t2 > 0.0 ? t1 + t2 : 0.0;
    double x2 = x[6] + cos_ang * tt_rate_n * dtinc;
    formR1(p, s, sn, x1, x2, R1);
    double nR1 = sqrt(R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] +
                      R1[3] * R1[3] + R1[4] * R1[4] + R1[5] * R1[5]);
    inR1 = nR1;
    bool ok = true;
    int iter = 0;
    while (nR1 > p.atol1 && nR1 / inR1 > p.rtol1) {
      double J11[36], dx1[6], wv[6];
      formJ11(p, s, x1, x2, J11);
      // J11^T R1 first: the factorisation below overwrites J11
#pragma unroll
      for (int j = 0; j < 6; ++j)
        wv[j] = J11[0 + 6 * j] * R1[0] + J11[1 + 6 * j] * R1[1] +
                J11[2 + 6 * j] * R1[2] + J11[3 + 6 * j] * R1[3] +
                J11[4 + 6 * j] * R1[4] + J11[5 + 6 * j] * R1[5];
      // J11 dx1 = -R1 in place: bitwise the solution of (-J11) dx1 = R1
      // (the same pivots and multipliers; every other value negated)
#pragma unroll
      for (int k = 0; k < 6; ++k) dx1[k] = -R1[k];
      lu_solve<6, 1>(J11, dx1);
      double alpha = 1.0;
      double dotR1 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] +
                     R1[3] * R1[3] + R1[4] * R1[4] + R1[5] * R1[5];
      double ls1 = 0.5 * dotR1;
      double ls2 = c * dot6(dx1, wv);
      int ls = 0;
      for (;;) {
        double nlsx = ls1 + ls2 * alpha, xn[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) xn[k] = x1[k] + alpha * dx1[k];
        formR1(p, s, sn, xn, x2, R1);
        dotR1 = R1[0] * R1[0] + R1[1] * R1[1] + R1[2] * R1[2] +
                R1[3] * R1[3] + R1[4] * R1[4] + R1[5] * R1[5];
        nR1 = sqrt(dotR1);
        double nRs = 0.5 * dotR1;
        if (nRs <= nlsx || ls > mls) {
#pragma unroll
          for (int k = 0; k < 6; ++k) x1[k] = xn[k];
          break;
        }
        alpha = red * alpha;
        ls = ls + 1;
      }
      iter = iter + 1;
      if (iter > p.miter || any_nan<6>(x1)) {
        fail = true;
        ok = false;
        break;
      }
    }
    if (ok) {
#pragma unroll
      for (int k = 0; k < 6; ++k) x[k] = x1[k];
      x[6] = x2;
    }
  }
  {  // corrector
    int iter = 0;
    double R[7];
    formR(p, s, sn, tau_n, x, R);
    double dotR = 0.0;
#pragma unroll
    for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
    double nR = sqrt(dotR), inR = nR;
    if (inR == 0.0) inR = inR1;
    while ((nR > p.atol && nR / inR > p.rtol) || iter < mmin) {
      double J[49], dx[7], wv[7];
      formJ(p, s, x, J);
      // the tangent takes the last iteration's Jacobian: parked in
      // global memory instead of held in registers across the loop
#pragma unroll
      for (int k = 0; k < 49; ++k) Jg[k * jstride] = J[k];
      dotR = 0.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
      double ls1 = 0.5 * dotR;
#pragma unroll
      for (int j = 0; j < 7; ++j) {
        double t = 0.0;
#pragma unroll
        for (int l = 0; l < 7; ++l) t = t + J[l + 7 * j] * R[l];
        wv[j] = t;
      }
      // J dx = -R in place: bitwise the solution of (-J) dx = R
#pragma unroll
      for (int k = 0; k < 7; ++k) dx[k] = -R[k];
      lu_solve<7, 1>(J, dx);
      double alpha = 1.0;
      double dd = 0.0;
#pragma unroll
      for (int k = 0; k < 7; ++k) dd = dd + dx[k] * wv[k];
      double ls2 = c * dd;
      int ls = 0;
      for (;;) {
        double nlsx = ls1 + ls2 * alpha, xn[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) xn[k] = x[k] + alpha * dx[k];
        formR(p, s, sn, tau_n, xn, R);
        dotR = 0.0;
#pragma unroll
        for (int k = 0; k < 7; ++k) dotR = dotR + R[k] * R[k];
        nR = sqrt(dotR);
        double nRs = 0.5 * dotR;
        if (nRs <= nlsx || ls > mls) {
#pragma unroll
          for (int k = 0; k < 7; ++k) x[k] = xn[k];
          break;
        }
        alpha = red * alpha;
        ls = ls + 1;
      }
      iter = iter + 1;
      if (iter > p.miter || any_nan<7>(x)) {
        fail = true;
        break;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) stress[k] = x[k];
  *tt = x[6];
}

// mm10_update_euler_angles (Kocks): angles of g (Rp R^T); with small
// strains (rot null, R = I) Rp R^T is Rp exactly
__device__ inline void euler_angles(const Props& p, const double* Rp_in,
                                    double* eul, const double* rot,
                                    int64_t SG) {
  const double pi = 3.141592653589793, one_eighty = 180.0;
  const double* g = p.cr->g;
  double Rp[9];
  if (rot) {  // mm10_a_mult_type_3t: Rp R^T
    double R[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) R[k] = rot[k * SG];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        Rp[i + 3 * j] = Rp_in[i] * R[j] + Rp_in[i + 3] * R[j + 3] +
                        Rp_in[i + 6] * R[j + 6];
  } else {
#pragma unroll
    for (int k = 0; k < 9; ++k) Rp[k] = Rp_in[k];
  }
  double full[9];
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      full[i + 3 * j] = g[i] * Rp[0 + 3 * j] + g[i + 3] * Rp[1 + 3 * j] +
                        g[i + 6] * Rp[2 + 3 * j];
  double psiK = atan2(full[2 + 3 * 1], full[2 + 3 * 0]);
  if (psiK < 0.0) psiK = psiK + 2.0 * pi;
  double phiK = atan2(full[1 + 3 * 2], full[0 + 3 * 2]);
  if (phiK < 0.0) phiK = phiK + 2.0 * pi;
  double f33 = full[8] > 1.0 ? 1.0 : full[8];
  double thetaK = acos(f33);
  if (p.angle_type == 1) {
    eul[0] = one_eighty / pi * psiK;
    eul[1] = one_eighty / pi * thetaK;
    eul[2] = one_eighty / pi * phiK;
  } else {
    eul[0] = psiK;
    eul[1] = thetaK;
    eul[2] = phiK;
  }
}

}  // namespace

// ---------------------------------------------------------------------
// One thread per Gauss point of an mm10 block (all its crystals).
// ---------------------------------------------------------------------
__device__ __forceinline__ void mm10_item(const DevModel& M,
                                          const WorkItem wi, int step,
                                          int iter0_off, double* s_geo) {
  if ((int)threadIdx.x >= wi.count) return;
  const BlockDesc bd = M.blocks[wi.blk];
  const MatDesc md = M.mats[bd.mat];
  const int G = bd.ngp;
  const int64_t SG = (int64_t)bd.span * G;
  const int64_t q = wi.first + threadIdx.x;  // point inside the block
  const int elem = (int)(q / G);
  const int ncry = bd.ncry;
  const double* hn = M.hist_n + bd.hist_off + q;
  double* hn1 = M.hist_n1 + bd.hist_off + q;
  const double* sn9 = M.stress_n + 9 * bd.q_off + q;
  double* sn1 = M.stress_n1 + 9 * bd.q_off + q;
  double* cp = M.cep + 21 * bd.q_off + q;
  const double* ud = M.uddt + 6 * bd.q_off + q;
  // finite strains: [R n+1] of the point (the recovery's), else R = I
  const double* rot = bd.geonl ? M.rot_tmp + 9 * bd.q_off + q : nullptr;
  // (inputs and sums are read from and kept in memory where they are
  // needed, not held in registers across the local solves)
  const double dt = M.dt;

  Props p;
  p.rate_n = md.cp.rate_n;
  p.tau_y = md.cp.tau_y;
  p.tau_v = md.cp.tau_v;
  p.voce_m = md.cp.voce_m;
  p.theta_0 = md.cp.theta_0;
  p.iD_v = md.cp.iD_v;
  p.eps_dot_0_y = md.cp.eps_dot_0_y;
  p.atol = md.cp.atol;
  p.atol1 = md.cp.atol1;
  p.rtol = md.cp.rtol;
  p.rtol1 = md.cp.rtol1;
  p.miter = md.cp.miter;
  p.alter_mode = md.cp.alter_mode;
  p.angle_type = md.cp.angle_type;

  // crystal sums of the tangent, the stresses and the slips accumulate in
  // their output words (history 1-36, stress n+1 1-6, history 76-87), read
  // back only by this thread
  double t_work = 0.0, p_work = 0.0, p_strain = 0.0;
  bool cut = false;

  for (int c = 0; c < ncry; ++c) {
    p.cr = M.crys + bd.crys_off + (int64_t)elem * ncry + c;
    const int64_t cb = (int64_t)(87 + 71 * c) * SG;
    // n-state of the crystal (mm10_copy_cc_hist; at step 1 the
    // mm10_init_cc_hist0 values)
    // (Dn, the n strain increment, is read where the predictor uses it)
    double sn[6], Rpn[9], tau_n, tt_rate_n;
    const double* Dn = step == 1 ? nullptr : hn + cb + 18 * SG;
    if (step == 1) {
#pragma unroll
      for (int k = 0; k < 6; ++k) sn[k] = sn9[k * SG];
#pragma unroll
      for (int k = 0; k < 9; ++k) Rpn[k] = (k % 4 == 0) ? 1.0 : 0.0;
      tau_n = p.tau_y + 1.0e-5;
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) sn[k] = hn[cb + k * SG];
#pragma unroll
      for (int k = 0; k < 9; ++k) Rpn[k] = hn[cb + (9 + k) * SG];
      tau_n = hn[cb + 42 * SG];
    }
    tt_rate_n = hn[cb + 58 * SG];

    // np1 (mm10_setup_np1 + mm10_setup)
    double* gl = s_geo + threadIdx.x;
    CState np1;
    slip_geometry(p, Rpn, gl, np1, rot, SG);
#pragma unroll
    for (int k = 0; k < 6; ++k) np1.D[k] = ud[k * SG];
    np1.tinc = dt;
    np1.tt_rate = 0.0;
    setup(p, np1);
    // mm10_solve_strup
    double stress[6], ostress[6], tt = tau_n, ott = tau_n;
#pragma unroll
    for (int k = 0; k < 6; ++k) stress[k] = ostress[k] = sn[k];
    double temp1 = dot6(np1.D, np1.D);
    double temp2 = dot6(stress, stress);
    const bool no_load = temp2 == 0.0 && temp1 == 0.0;
    // the last corrector Jacobian lives in the crystal's n+1 history words
    // 0-48 until the tangent has read it (mm10_store_cryhist then writes
    // those words)
    double* Jg = hn1 + cb;
    bool fail = false;
    double curr_tt_rate = 0.0;
    double tangent[36];
    // the crystal's tangent into the point's average (history words 0-35;
    // stored as soon as it is formed, so that it does not stay live)
    auto store_tangent = [&]() {
      if (c == 0) {
#pragma unroll
        for (int k = 0; k < 36; ++k) hn1[k * SG] = tangent[k];
      } else {
#pragma unroll
        for (int k = 0; k < 36; ++k) hn1[k * SG] = hn1[k * SG] + tangent[k];
      }
    };
    if (iter0_off || no_load) {
      const double* C = p.cr->C;
#pragma unroll
      for (int k = 0; k < 36; ++k) tangent[k] = C[k];
      store_tangent();
      if (!no_load) {
        double R1[6];
        formR1(p, np1, sn, stress, tt, R1);
#pragma unroll
        for (int k = 0; k < 6; ++k) stress[k] = stress[k] - R1[k];
      }
    } else {
      // the sub-steps run on np1 itself, their increments scaled from
      // the point's full-step values in memory (ud, dt); np1's full-step
      // D, tinc and dg are formed again after the loop, so that they are
      // not held across the local solves
      double frac = 0.0, sstep = 1.0;
      int cuts = 0;
      CState& curr = np1;
      while (frac < 1.0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) curr.D[k] = ud[k * SG] * (sstep + frac);
        curr.tinc = dt * (sstep + frac);
        curr.tt_rate = 0.0;
        setup(p, curr);
        tt = tau_n;
        solve(p, curr, sn, tau_n, tt_rate_n, Dn, SG, stress, &tt, fail,
              dt * sstep, Jg, SG);
        if (fail) {
#pragma unroll
          for (int k = 0; k < 6; ++k) stress[k] = ostress[k];
          tt = ott;
          sstep = sstep * 0.5;
          cuts = cuts + 1;
          if (cuts > 4) break;
          fail = false;
        } else {
#pragma unroll
          for (int k = 0; k < 6; ++k) ostress[k] = stress[k];
          ott = tt;
          frac = frac + sstep;
        }
      }
      if (fail || isnan(tt) || any_nan<6>(stress)) {
        cut = true;
        break;
      }
      curr_tt_rate = curr.tt_rate;
#pragma unroll
      for (int k = 0; k < 6; ++k) np1.D[k] = ud[k * SG];
      np1.tinc = dt;
      setup(p, np1);
    }
    np1.tt_rate = curr_tt_rate;
    double eul[3] = {0, 0, 0}, Rp1[9], ee[6], slip[NS], u[15], ep6[6],
           ed6[6];
    double work_inc = 0.0, p_work_inc = 0.0, p_strain_inc = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) Rp1[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) ee[k] = ep6[k] = ed6[k] = 0.0;
#pragma unroll
    for (int k = 0; k < NS; ++k) slip[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 15; ++k) u[k] = 0.0;
    if (!(iter0_off || no_load)) {
      // ---- mm10_tangent: T = (J11 - J12 J21 / J22)^-1 C, symmetrised
      // unless the assembly is asymmetric
      double J[49], JJ[36], beta[12];
#pragma unroll
      for (int k = 0; k < 49; ++k) J[k] = Jg[k * SG];
#pragma unroll
      for (int cc = 0; cc < 6; ++cc)
#pragma unroll
        for (int r = 0; r < 6; ++r) JJ[r + 6 * cc] = J[r + 7 * cc];
      double a1 = J[48];
#pragma unroll
      for (int cc = 0; cc < 6; ++cc) {
        beta[cc] = J[6 + 7 * cc];
        beta[6 + cc] = 0.0;
      }
      lu_solve<1, 12>(&a1, beta);
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double t = -1.0 * beta[j];
        if (t != 0.0) {
#pragma unroll
          for (int i = 0; i < 6; ++i) JJ[i + 6 * j] = JJ[i + 6 * j] + t * J[i + 42];
        }
      }
      const double* C = p.cr->C;
#pragma unroll
      for (int k = 0; k < 36; ++k) tangent[k] = C[k] - 0.0 - 0.0;
      lu_solve<6, 6>(JJ, tangent);
      if (!M.asym) {  // kept as formed for asymmetric assembly (:1165)
        double w[36];
#pragma unroll
        for (int j = 0; j < 6; ++j)
#pragma unroll
          for (int i = 0; i < 6; ++i)
            w[i + 6 * j] = (tangent[i + 6 * j] + tangent[j + 6 * i]) * 0.5;
#pragma unroll
        for (int k = 0; k < 36; ++k) tangent[k] = w[k];
      }
      store_tangent();
      // ---- slips at the converged state (np1: full-step dg)
      double rs[NS], gam[NS];
      slips(p, np1, stress, tt, rs, gam);
      // ---- mm10_update_rotation: Rp = expm(W(wbarp)) Rp_n (Rp_n read
      // again here)
      {
        double Rpe[9];
#pragma unroll
        for (int k = 0; k < 9; ++k)
          Rpe[k] = step == 1 ? ((k % 4 == 0) ? 1.0 : 0.0)
                             : hn[cb + (9 + k) * SG];
        double wb[3];
        if (rot) {  // form_wbarp takes qs, of Rp_n alone
          CState nq = np1;
          double tn[9];
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) tn[i + 3 * j] = Rpe[j + 3 * i];
          rt2rvw(tn, nq.rw);
          form_w(p, nq, rs, gam, wb);
        } else {
          form_w(p, np1, rs, gam, wb);
        }
        double W[9] = {0.0, -wb[2], -wb[1], wb[2], 0.0, -wb[0],
                       wb[1], wb[0], 0.0};  // column-major of wv2wt
        double alpha = sqrt(W[1 + 3 * 2] * W[1 + 3 * 2] +
                            W[0 + 3 * 2] * W[0 + 3 * 2] +
                            W[0 + 3 * 1] * W[0 + 3 * 1]);
        double A[9];
        if (alpha < 1.0e-16) {
#pragma unroll
          for (int k = 0; k < 9; ++k) A[k] = 0.0;
        } else {
          double ca = (1.0 - cos(alpha)) / (alpha * alpha);
          double sa = sin(alpha) / alpha;
#pragma unroll
          for (int j = 0; j < 3; ++j) {
#pragma unroll
            for (int i = 0; i < 3; ++i) A[i + 3 * j] = sa * W[i + 3 * j];
#pragma unroll
            for (int l = 0; l < 3; ++l) {
              double t = ca * W[l + 3 * j];
#pragma unroll
              for (int i = 0; i < 3; ++i)
                A[i + 3 * j] = A[i + 3 * j] + t * W[i + 3 * l];
            }
          }
        }
        A[0] = A[0] + 1.0;
        A[4] = A[4] + 1.0;
        A[8] = A[8] + 1.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
          for (int i = 0; i < 3; ++i)
            Rp1[i + 3 * j] = A[i] * Rpe[0 + 3 * j] + A[i + 3] * Rpe[1 + 3 * j] +
                             A[i + 6] * Rpe[2 + 3 * j];
      }
      // ---- mm10_output
#pragma unroll
      for (int i = 0; i < NS; ++i) slip[i] = gam[i];
      euler_angles(p, Rp1, eul, rot, SG);
      double dif[NS], t1, t2;
#pragma unroll
      for (int i = 0; i < NS; ++i) dif[i] = rs[i] * np1.tinc * p.iD_v;
#pragma unroll
      for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int k = 0; k < 6; ++k) ed6[k] = ed6[k] + dif[i] * np1.ms(i, k);
      double edt[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) edt[k] = ed6[k];
#pragma unroll
      for (int k = 0; k < 6; ++k) ed6[k] = edt[k] / np1.tinc;
      t1 = edt[0] * edt[0] + edt[1] * edt[1] + edt[2] * edt[2];
      t2 = edt[3] * edt[3] + edt[4] * edt[4] + edt[5] * edt[5];
      double ed_dot = sqrt(2.0 / 3.0 * (t1 + 0.5 * t2));
      u[14] = ed_dot / np1.tinc;
      work_inc = dot6(stress, np1.D);
#pragma unroll
      for (int k = 0; k < 6; ++k) ee[k] = stress[k];
      chol_solve6(p.cr->C, ee);
      if (rot) {  // ee = RT2RVE(R) C^-1 sigma (exact for R = I)
        double R[9], er[36], t[6];
#pragma unroll
        for (int k = 0; k < 9; ++k) R[k] = rot[k * SG];
        rt2rve(R, er);
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          double a = er[i] * ee[0];
#pragma unroll
          for (int j = 1; j < 6; ++j) a = a + er[i + 6 * j] * ee[j];
          t[i] = a;
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) ee[k] = t[k];
      }
      double dbarp[6] = {0, 0, 0, 0, 0, 0}, wp[3], ewwe[6], ep[6];
#pragma unroll
      for (int i = 0; i < NS; ++i)
#pragma unroll
        for (int k = 0; k < 6; ++k) dbarp[k] = dbarp[k] + slip[i] * np1.ms(i, k);
      form_w(p, np1, rs, gam, wp);
      symsw(ee, wp, ewwe);
#pragma unroll
      for (int k = 0; k < 6; ++k) ep[k] = dbarp[k] + ewwe[k];
#pragma unroll
      for (int k = 0; k < 6; ++k) ep6[k] = ep[k] / np1.tinc;
      t1 = ep[0] * ep[0] + ep[1] * ep[1] + ep[2] * ep[2];
      t2 = ep[3] * ep[3] + ep[4] * ep[4] + ep[5] * ep[5];
      u[10] = sqrt(2.0 / 3.0 * (t1 + 0.5 * t2)) / np1.tinc;
#pragma unroll
      for (int k = 0; k < 6; ++k) ep[k] = ep[k] + edt[k];
      t1 = ep[0] * ep[0] + ep[1] * ep[1] + ep[2] * ep[2];
      t2 = ep[3] * ep[3] + ep[4] * ep[4] + ep[5] * ep[5];
      p_strain_inc = sqrt(2.0 / 3.0 * (t1 + 0.5 * t2));
      p_work_inc = dot6(stress, ep);
      double ec_dot = p_strain_inc / np1.tinc, n_eff;
      if (ec_dot > 0.0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) ep[k] = ep[k] / np1.tinc;
        n_eff = 0.0;
        const double ptt = vpow(tt, p.rate_n);
#pragma unroll
        for (int i = 0; i < NS; ++i) {
          double d = vpow(fabs(rs[i]), p.rate_n - 1.0);
          d = np1.dg * p.rate_n / ptt * d;
          double dgdt = d + np1.tinc * p.iD_v;
          double a = np1.ms(i, 0) * ep[0] + np1.ms(i, 1) * ep[1] +
                     np1.ms(i, 2) * ep[2];
          double b = np1.ms(i, 3) * ep[3] + np1.ms(i, 4) * ep[4] +
                     np1.ms(i, 5) * ep[5];
          double ec_slip = a + 0.5 * b;
          double q1 = rs[i] * dgdt / ec_dot;
          double q2 = ec_slip / ec_dot / np1.tinc;
          n_eff = n_eff + (2.0 / 3.0) * q1 * q2;
        }
      } else {
        n_eff = 1.0e10;
      }
      u[11] = n_eff;
      double s_tr = (stress[0] + stress[1] + stress[2]) / 3.0;
      double se0 = stress[0] - s_tr, se1 = stress[1] - s_tr,
             se2 = stress[2] - s_tr;
      t1 = se0 * se0 + se1 * se1 + se2 * se2;
      t2 = stress[3] * stress[3] + stress[4] * stress[4] +
           stress[5] * stress[5];
      s_tr = sqrt(1.5 * (t1 + 2.0 * t2));
      u[12] = s_tr;
      double B_eff;
      if (ec_dot < 1.0e-100)
        B_eff = 0.0;
      else if (n_eff > 100.0)
        B_eff = -1.0;
      else
        B_eff = ec_dot / pow(s_tr, n_eff);
      u[13] = B_eff;
#pragma unroll
      for (int i = 0; i < NS; ++i) slip[i] = slip[i] + dif[i];
      double maxslip = 0.0;
      int sysID = 0;
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        double cs = fabs(slip[i]);
        if (cs > maxslip) {
          maxslip = cs;
          sysID = i + 1;
        }
      }
      u[5] = maxslip / np1.tinc;
      u[6] = (double)sysID;
      int numAct = 0;
      maxslip = 0.1 * maxslip;
#pragma unroll
      for (int i = 0; i < NS; ++i)
        if (fabs(slip[i]) >= maxslip) numAct = numAct + 1;
      u[7] = (double)numAct;
    }
    // accumulate for the averages
#pragma unroll
    for (int k = 0; k < 6; ++k)
      sn1[k * SG] = c == 0 ? stress[k] : sn1[k * SG] + stress[k];
#pragma unroll
    for (int k = 0; k < NS; ++k)
      hn1[(75 + k) * SG] = c == 0 ? slip[k] : hn1[(75 + k) * SG] + slip[k];
    t_work = t_work + work_inc;
    p_work = p_work + p_work_inc;
    p_strain = p_strain + p_strain_inc;
    // mm10_store_cryhist
    double* h1 = hn1 + cb;
#pragma unroll
    for (int k = 0; k < 6; ++k) h1[k * SG] = stress[k];
#pragma unroll
    for (int k = 0; k < 3; ++k) h1[(6 + k) * SG] = eul[k];
#pragma unroll
    for (int k = 0; k < 9; ++k) h1[(9 + k) * SG] = Rp1[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) h1[(18 + k) * SG] = np1.D[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) h1[(24 + k) * SG] = ee[k];
#pragma unroll
    for (int k = 0; k < NS; ++k) h1[(30 + k) * SG] = slip[k];
    h1[42 * SG] = tt;
#pragma unroll
    for (int k = 0; k < 15; ++k) h1[(43 + k) * SG] = u[k];
    h1[58 * SG] = np1.tt_rate;
#pragma unroll
    for (int k = 0; k < 6; ++k) h1[(59 + k) * SG] = ep6[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) h1[(65 + k) * SG] = ed6[k];
  }
  if (cut) {
    atomicOr(&M.flags[1], 1);
    return;
  }
  const double rn = (double)ncry;
  double tang_avg[36];
#pragma unroll
  for (int k = 0; k < 36; ++k) tang_avg[k] = hn1[k * SG] / rn;
  t_work = t_work / rn;
  p_work = p_work / rn;
  p_strain = p_strain / rn;
  // mm10_a_store_crystal
#pragma unroll
  for (int k = 0; k < 6; ++k) sn1[k * SG] = sn1[k * SG] / rn;
  sn1[6 * SG] = sn9[6 * SG] + t_work;
  sn1[7 * SG] = sn9[7 * SG] + p_work;
  sn1[8 * SG] = sn9[8 * SG] + p_strain;
#pragma unroll
  for (int k = 0; k < 36; ++k) hn1[k * SG] = tang_avg[k];
#pragma unroll
  for (int k = 36; k < 63; ++k) hn1[k * SG] = 0.0;
#pragma unroll
  for (int k = 0; k < 9; ++k)
    hn1[(63 + k) * SG] = rot ? rot[k * SG] : ((k % 4 == 0) ? 1.0 : 0.0);
  double w72 = step == 1 ? 0.0 : hn[72 * SG];
  double w73 = step == 1 ? 0.0 : hn[73 * SG];
  double w74 = step == 1 ? 0.0 : hn[74 * SG];
  hn1[72 * SG] = w72 + t_work;
  hn1[73 * SG] = w73 + p_work;
  hn1[74 * SG] = w74 + p_strain;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    double base = step == 1 ? 0.0 : hn[(75 + k) * SG];
    hn1[(75 + k) * SG] = base + hn1[(75 + k) * SG] / rn;
  }
  // [D] for the tangent kernel: the lower triangle of history words 1-36
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int cc = 0; cc <= r; ++cc)
      cp[tri(r, cc) * SG] = tang_avg[r + 6 * cc];
}

// ---------------------------------------------------------------------
// grad Fe of mm10 blocks: rknstr_finish_cp (rknstr.f) with
// mm10_calc_grads (mm10_a.f:1579).  Small strains: for single-crystal
// elements the
// reference fits each component of Rt3d(3,3,ngp) linearly over the Gauss
// points (DGELS) and stores the three slopes of every component, the same
// 27 words at every point; Rt3d is a flat copy (mm10_a_copy_vector) of
// 9 ngp consecutive words of the block's gathered n-history
// elem_hist(span, hist, ngp) from (element 1, first Rp word of crystal 1,
// point 1), so words of several elements of the block enter it.  At step
// 1 those words are the values mm10 initialised in place.  Elements with
// several crystals get zeros.  Finite strains: per element (below).  One
// 64-lane workgroup per mm10 block.
// ---------------------------------------------------------------------
__device__ double mm10_init_word(const DevModel& M, const BlockDesc& bd,
                                 int w, int i, int64_t q, int64_t SG) {
  const double* hn = M.hist_n + bd.hist_off;
  if (w < kMM10Common)  // mm10_init_general_hist, _uout_hist, _slip_hist
    return (w >= 63 && w < 72) ? ((w - 63) % 4 == 0 ? 1.0 : 0.0) : 0.0;
  const int c = (w - kMM10Common) / kMM10Crystal;
  const int k = (w - kMM10Common) % kMM10Crystal;
  if (k < 6) return M.stress_n[9 * bd.q_off + k * SG + q];  // stress
  if (k < 9)                                                // angles
    return M.crys[bd.crys_off + (int64_t)i * bd.ncry + c].angles[k - 6];
  if (k < 18) return (k - 9) % 4 == 0 ? 1.0 : 0.0;          // Rp
  if (k == 42) return M.mats[bd.mat].cp.tau_y + 1.0e-5;     // tau tilde
  if (k < 58) return 0.0;
  return hn[w * SG + q];  // words mm10 does not initialise
}

__global__ void __launch_bounds__(64)
    k_mm10_grads(DevModel M, const int32_t* __restrict__ blks, int lo,
                 int hi, int step) {
  const int b = blks[blockIdx.x];
  if (b < lo || b >= hi) return;
  const BlockDesc bd = M.blocks[b];
  const GaussTable* gt = M.tables + bd.table;
  const int S = bd.span, G = bd.ngp, H = bd.hist;
  const int64_t SG = (int64_t)S * G;
  const int l = threadIdx.x;
  __shared__ double s_v[9 * kMaxGp];
  __shared__ double s_g[27];
  const double* hn = M.hist_n + bd.hist_off;
  if (bd.ncry == 1) {
    for (int f = l; f < 9 * G; f += 64) {
      const int64_t L = (int64_t)(kMM10Common + 9) * S + f;
      const int i = (int)(L % S);
      const int w = (int)((L / S) % H);
      const int g = (int)(L / ((int64_t)S * H));
      const int64_t q = (int64_t)i * G + g;
      s_v[f] = step == 1 ? mm10_init_word(M, bd, w, i, q, SG) : hn[w * SG + q];
    }
  }
  __syncthreads();
  if (l < 27) {
    // grads(a, b, d) at a + 3b + 9d from Rt3d(a, b, p) at a + 3b + 9p
    double x = 0.0;
    if (bd.ncry == 1) {
      const int ab = l % 9, d = l / 9;
      for (int p = 0; p < G; ++p) x = x + gt->lsq[d][p] * s_v[ab + 9 * p];
    }
    s_g[l] = x;
  }
  __syncthreads();
  double* hn1 = M.hist_n1 + bd.hist_off;
  if (!bd.geonl) {
    for (int64_t k = l; k < 27 * SG; k += 64)
      hn1[(36 + k / SG) * SG + k % SG] = s_g[k / SG];
    return;
  }
  // finite strains (mm10_calc_grads' geonl branch): per element, the fit
  // of Rp R^T (Rp of crystal 1 from the element's own n-history, R = [R
  // n+1]) taken to the current axes by the inverse (mm10_a_invert_33) of
  // the element's n = 0 Jacobian at its last point; zeros for several
  // crystals or a singular Jacobian
  for (int i = l; i < S; i += 64) {
    double gv[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) gv[k] = 0.0;
    const double* a = M.jac_tmp + 9 * ((int64_t)bd.felem - 1 + i);
#define A3(r, c) a[3 * ((r)-1) + (c)-1]
    const double det =
        A3(1, 1) * A3(2, 2) * A3(3, 3) - A3(1, 1) * A3(2, 3) * A3(3, 2) -
        A3(1, 2) * A3(2, 1) * A3(3, 3) + A3(1, 2) * A3(2, 3) * A3(3, 1) +
        A3(1, 3) * A3(2, 1) * A3(3, 2) - A3(1, 3) * A3(2, 2) * A3(3, 1);
    if (bd.ncry == 1 && fabs(det) > 1.0e-10) {
      double ji[9];  // jacinv row-major: transpose(cofactor) / det
      ji[0] = (A3(2, 2) * A3(3, 3) - A3(2, 3) * A3(3, 2)) / det;
      ji[3] = -(A3(2, 1) * A3(3, 3) - A3(2, 3) * A3(3, 1)) / det;
      ji[6] = (A3(2, 1) * A3(3, 2) - A3(2, 2) * A3(3, 1)) / det;
      ji[1] = -(A3(1, 2) * A3(3, 3) - A3(1, 3) * A3(3, 2)) / det;
      ji[4] = (A3(1, 1) * A3(3, 3) - A3(1, 3) * A3(3, 1)) / det;
      ji[7] = -(A3(1, 1) * A3(3, 2) - A3(1, 2) * A3(3, 1)) / det;
      ji[2] = (A3(1, 2) * A3(2, 3) - A3(1, 3) * A3(2, 2)) / det;
      ji[5] = -(A3(1, 1) * A3(2, 3) - A3(1, 3) * A3(2, 1)) / det;
      ji[8] = (A3(1, 1) * A3(2, 2) - A3(1, 2) * A3(2, 1)) / det;
#undef A3
      double sl[27];  // slopes (d) of component (a, b) at a + 3 b + 9 d
#pragma unroll
      for (int k = 0; k < 27; ++k) sl[k] = 0.0;
      for (int g = 0; g < G; ++g) {
        const int64_t q = (int64_t)i * G + g;
        double Rp[9], R[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          const int w = kMM10Common + 9 + k;  // Rp of crystal 1
          Rp[k] = step == 1 ? mm10_init_word(M, bd, w, i, q, SG)
                            : hn[w * SG + q];
          R[k] = M.rot_tmp[9 * bd.q_off + k * SG + q];
        }
#pragma unroll
        for (int ia = 0; ia < 3; ++ia)
#pragma unroll
          for (int ib = 0; ib < 3; ++ib) {
            // mm10_a_mult_type_3t: (Rp R^T)(ia, ib)
            const double v = Rp[ia] * R[ib] + Rp[ia + 3] * R[ib + 3] +
                             Rp[ia + 6] * R[ib + 6];
#pragma unroll
            for (int d = 0; d < 3; ++d)
              sl[ia + 3 * ib + 9 * d] = sl[ia + 3 * ib + 9 * d] +
                                        gt->lsq[d][g] * v;
          }
      }
#pragma unroll
      for (int ab = 0; ab < 9; ++ab)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          gv[ab + 9 * d] = ji[3 * d] * sl[ab] + ji[3 * d + 1] * sl[ab + 9] +
                           ji[3 * d + 2] * sl[ab + 18];
    }
    for (int g = 0; g < G; ++g) {
      const int64_t q = (int64_t)i * G + g;
#pragma unroll
      for (int k = 0; k < 27; ++k) hn1[(36 + k) * SG + q] = gv[k];
    }
  }
}

int launch_mm10_grads(const DevModel& M, const int32_t* d_blks, int nblks,
                      int lo, int hi, int step, hipStream_t s) {
  if (nblks > 0)
    hipLaunchKernelGGL(k_mm10_grads, dim3(nblks), dim3(64), 0, s, M, d_blks,
                       lo, hi, step);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

// Persistent over the work list: each workgroup (one wave) takes the next
// item of 64 points from a device counter until the list is exhausted, so
// points whose local solves sub-step or iterate longer do not hold back a
// fixed share of the grid.  The counter starts at zero (launch_mm10 clears
// it with a fill kernel first).
__global__ void __launch_bounds__(64)
    k_mm10(DevModel M, const WorkItem* __restrict__ work, int nwork,
           int* __restrict__ counter, int step, int iter0_off) {
  __shared__ double s_geo[kGeoWords * kLanes];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= nwork) break;  // uniform across the workgroup
    mm10_item(M, work[item], step, iter0_off, s_geo);
  }
}

// The batched form: the items of every slot in one list, slot s owning
// global items [item0_s, item0_s + nwork_s); the workgroup finds its slot
// by bisection over the slots' item0 (ascending).
__global__ void __launch_bounds__(64)
    k_mm10_batch(const Mm10Slot* __restrict__ slots, int nslot,
                 int nwork_total, int* __restrict__ counter) {
  __shared__ double s_geo[kGeoWords * kLanes];
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (item >= nwork_total) break;  // uniform across the workgroup
    int lo = 0, hi = nslot - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (slots[mid].item0 <= item)
        lo = mid;
      else
        hi = mid - 1;
    }
    const Mm10Slot& sl = slots[lo];
    mm10_item(sl.M, sl.work[item - sl.item0], sl.step, sl.iter0_off, s_geo);
  }
}

int launch_mm10_batch(const Mm10Slot* d_slots, int nslot, int nwork_total,
                      int* d_counter, int grid, hipStream_t s) {
  if (nwork_total <= 0 || nslot <= 0) return W3D_OK;
  int rc = launch_fill_i32(d_counter, 1, 0, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mm10_batch,
                     dim3(grid < nwork_total ? grid : nwork_total), dim3(64),
                     0, s, d_slots, nslot, nwork_total, d_counter);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_mm10(const DevModel& M, const WorkItem* d_work, int nwork,
                int* d_counter, int grid, int step, int iter0_off,
                hipStream_t s) {
  if (nwork <= 0) return W3D_OK;
  int rc = launch_fill_i32(d_counter, 1, 0, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mm10, dim3(grid < nwork ? grid : nwork), dim3(64), 0,
                     s, M, d_work, nwork, d_counter, step, iter0_off);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

}  // namespace w3d
