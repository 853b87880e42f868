// w3d_kernels.hip -- gfx950 kernels of the element-block hot path:
// strain/stress recovery + mm01 + element internal forces, the deterministic
// gathers that assemble element forces and block tangents, and the FP64
// vector form of the block tangent.
//
// Reference routines restated (file:line under the reference's src/):
//   rknstr.f:563 rknstr_sm_displ, gtmat1.f:300 jacob1, vol_terms.f:14,
//   vol_avg.f:13, blcmp1.f:15, bmod.f:16, gtlsn1.f:25, rstgp1.f:351 rstgp2,
//   rstgp1.f:614 drive_01_update (+ :998 drive_01_update_b), mm01.f:24-1348
//   (mm01_set_history, mm01_init, mm01_general, mm01_sig_final, mm01_wrapup,
//   cnst1), rstgp1.f:5360 rstgp1_store_cep, rstgp1.f:5189 rstgp1_b,
//   gpifv1.f:16, rknifv.f:141, addifv.f:16, gptns1.f:1174
//   extract_D_symmetric, gptns1.f:1628 bdbtgen, gpmas1.f:399 adjust_cnst,
//   assemble_code.f:659 assem_by_row.
#include <hip/hip_runtime.h>

#include "w3d_internal.h"

namespace w3d {

// ---------------------------------------------------------------------
// mm01 (bilinear Mises, isothermal) + cnst1 for one Gauss point.
// cgn: 9 stress words at n (words 8, 9 may be reset at step 1);
// h: 11 history words at n; outputs cgn1 (9), h1 (11), cep (21 lower
// triangle, row-major (1,1),(2,1),(2,2),...).  Returns false on a negative
// plastic multiplier (die_abort in mm01_general).
// ---------------------------------------------------------------------
__device__ inline double state_word(int st) {
  return __longlong_as_double((long long)(unsigned)st);
}
__device__ inline int state_of(double w) {
  return (int)(__double_as_longlong(w) & 0xffffffffll);
}

__device__ bool mm01_point(int step, const MatDesc& md, double* cgn,
                           double* cgn1, const double* deps, double* h,
                           double* h1, double first_hist11, double* cep) {
  const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, half = 0.5,
               six = 6.0, four = 4.0;
  const double e = md.e, nu = md.nu, beta = md.beta, hprime = md.hprime,
               yld = md.yld;
  const double root3 = sqrt(3.0), root2 = sqrt(2.0);
  if (step == 1) {  // mm01_set_history
    double kn = yld / root3;
    h[0] = zero; h[1] = kn; h[2] = zero; h[3] = state_word(3);
    h[4] = hprime;
#pragma unroll
    for (int k = 5; k < 11; ++k) h[k] = zero;
    cgn[7] = zero;
    cgn[8] = zero;
    first_hist11 = zero;
  }
  // mm01_init
  const double htol = 0.000001, yld_tol = 0.0000001;
  double deps_vol = deps[0] + deps[1] + deps[2];
  double eps_mean = deps_vol / three;
  double de1 = deps[0] - eps_mean, de2 = deps[1] - eps_mean,
         de3 = deps[2] - eps_mean;
  double g_n = e / two / (one + nu);
  double een1 = (cgn[0] - nu * (cgn[1] + cgn[2])) / e;
  double een2 = (cgn[1] - nu * (cgn[0] + cgn[2])) / e;
  double een3 = (cgn[2] - nu * (cgn[0] + cgn[1])) / e;
  double een4 = cgn[3] / g_n, een5 = cgn[4] / g_n, een6 = cgn[5] / g_n;
  double eps_vol_n1 = een1 + een2 + een3 + deps_vol;
  double eps_mean_n = (een1 + een2 + een3) / three;
  double shear_mod = e / (two * (one + nu));
  double dse[6];
  dse[0] = two * shear_mod * ((een1 - eps_mean_n) + de1);
  dse[1] = two * shear_mod * ((een2 - eps_mean_n) + de2);
  dse[2] = two * shear_mod * ((een3 - eps_mean_n) + de3);
  dse[3] = shear_mod * (een4 + deps[3]);
  dse[4] = shear_mod * (een5 + deps[4]);
  dse[5] = shear_mod * (een6 + deps[5]);
  double hbari_np1 = beta * hprime;
  double hbark_np1 = (one - beta) * hprime;
  double hbark_n = (one - beta) * h[4];
  double kbar = (yld + hbari_np1 * h[2]) / root3;
  double lk = one;
  if (fabs(hbark_n) > htol) lk = hbark_np1 / hbark_n;
  double r[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) r[k] = dse[k] - h[5 + k] * lk;
  double mrts = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] +
                     two * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]));
  double yf = mrts - root2 * kbar;
#pragma unroll
  for (int k = 0; k < 6; ++k) r[k] = dse[k] - h[5 + k];
  mrts = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] +
              two * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]));
  bool yield = !md.lnelas && yf >= yld_tol * root2 * kbar;
  double dev[6];
  bool ok = true;
  if (yield) {  // mm01_general
    const double r2 = 1.414213562373095, twthrd = 0.666666666666667,
                 root23 = 0.816496580927, root2o3 = 0.4714045207910;
    double a = (two * shear_mod + twthrd * hbark_np1) / mrts;
    double b = one - lk;
    double d = root2o3 * hbari_np1;
    double vbar = mrts * mrts;
    double wbar = h[5] * h[5] + h[6] * h[6] + h[7] * h[7] +
                  two * (h[8] * h[8] + h[9] * h[9] +
                         first_hist11 * first_hist11);
    double vwbar = r[0] * h[5] + r[1] * h[6] + r[2] * h[7] +
                   two * (r[3] * h[8] + r[4] * h[9] + r[5] * h[10]);
    double t1 = a * a * vbar - two * d * d;
    double t2 = -four * d * kbar - two * a * vbar - two * a * b * vwbar;
    double t3 = b * b * wbar + two * b * vwbar + vbar - two * kbar * kbar;
    double discr = t2 * t2 - four * t1 * t3;
    double ldt = zero;
    if (discr >= zero) {
      double sq = sqrt(discr);
      double q1 = (-t2 + sq) / (two * t1);
      double q2 = (-t2 - sq) / (two * t1);
      ldt = q1 < q2 ? q1 : q2;
    }
    if (ldt < zero) ok = false;
    h1[0] = ldt;
    h1[1] = kbar + (r2 / three) * hbari_np1 * ldt;
    h1[2] = h[2] + ldt * root23;
    h1[4] = hprime;
    double const1 = twthrd * hbark_np1 * ldt / mrts;
    double const2 = one - ldt * (twthrd * hbark_np1 + two * shear_mod) / mrts;
#pragma unroll
    for (int k = 0; k < 6; ++k) h1[5 + k] = lk * h[5 + k] + const1 * r[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) dev[k] = h1[5 + k] + b * h[5 + k] + const2 * r[k];
  } else {
    h1[0] = zero;
    h1[1] = md.lnelas ? zero : kbar;
    h1[2] = h[2];
    h1[4] = hprime;
#pragma unroll
    for (int k = 0; k < 6; ++k) h1[5 + k] = h[5 + k] * lk;
#pragma unroll
    for (int k = 0; k < 6; ++k) dev[k] = r[k] + h[5 + k];
  }
  // mm01_sig_final
  double sig_mean =
      eps_vol_n1 *
      (three * e * nu / ((one + nu) * (one - two * nu)) + two * shear_mod) /
      three;
  cgn1[0] = dev[0] + sig_mean;
  cgn1[1] = dev[1] + sig_mean;
  cgn1[2] = dev[2] + sig_mean;
  cgn1[3] = dev[3];
  cgn1[4] = dev[4];
  cgn1[5] = dev[5];
  cgn1[6] = cgn[6] + half * (deps[0] * (cgn1[0] + cgn[0]) +
                             deps[1] * (cgn1[1] + cgn[1]) +
                             deps[2] * (cgn1[2] + cgn[2]) +
                             deps[3] * (cgn1[3] + cgn[3]) +
                             deps[4] * (cgn1[4] + cgn[4]) +
                             deps[5] * (cgn1[5] + cgn[5]));
  h1[3] = state_word(yield ? 1 : 3);
  // mm01_wrapup
  cgn1[7] = cgn[7];
  cgn1[8] = cgn[8];
  if (yield) {
    const double* s = cgn;
    double sbn = sqrt((s[0] - s[1]) * (s[0] - s[1]) +
                      (s[1] - s[2]) * (s[1] - s[2]) +
                      (s[0] - s[2]) * (s[0] - s[2]) +
                      six * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])) /
                 root2;
    s = cgn1;
    double sbn1 = sqrt((s[0] - s[1]) * (s[0] - s[1]) +
                       (s[1] - s[2]) * (s[1] - s[2]) +
                       (s[0] - s[2]) * (s[0] - s[2]) +
                       six * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])) /
                  root2;
    double depbar = h1[2] - h[2];
    cgn1[7] = cgn[7] + half * (sbn1 + sbn) * depbar;
    cgn1[8] = cgn[8] + depbar;
  }
  // cnst1, written straight into the stored lower triangle
  if (!yield) {
    double c1 = (e / ((one + nu) * (one - two * nu)));
    double c2 = (one - nu) * c1;
    double c3 = ((one - two * nu) / two) * c1;
    double c4 = nu * c1;
    cep[0] = c2;
    cep[1] = c4; cep[2] = c2;
    cep[3] = c4; cep[4] = c4; cep[5] = c2;
    cep[6] = zero; cep[7] = zero; cep[8] = zero; cep[9] = c3;
    cep[10] = zero; cep[11] = zero; cep[12] = zero; cep[13] = zero;
    cep[14] = c3;
    cep[15] = zero; cep[16] = zero; cep[17] = zero; cep[18] = zero;
    cep[19] = zero; cep[20] = c3;
  } else {
    const double root2c = 1.414213562;
    double kn1 = h1[1], hp = h1[4], ldt = h1[0];
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
    cep[0] = (albar + two * gbar - thbar * (r[0] * r[0]) / mrtsq);
    cep[1] = (albar - thbar * r[0] * r[1] / mrtsq);
    cep[2] = (albar + two * gbar - thbar * (r[1] * r[1]) / mrtsq);
    cep[3] = (albar - thbar * r[0] * r[2] / mrtsq);
    cep[4] = (albar - thbar * r[2] * r[1] / mrtsq);
    cep[5] = (albar + two * gbar - thbar * (r[2] * r[2]) / mrtsq);
    cep[6] = -(thbar * r[0] * r[3] / mrtsq);
    cep[7] = -(thbar * r[1] * r[3] / mrtsq);
    cep[8] = -(thbar * r[2] * r[3] / mrtsq);
    cep[9] = (gbar - thbar * (r[3] * r[3]) / mrtsq);
    cep[10] = -(thbar * r[0] * r[4] / mrtsq);
    cep[11] = -(thbar * r[1] * r[4] / mrtsq);
    cep[12] = -(thbar * r[2] * r[4] / mrtsq);
    cep[13] = -(thbar * r[3] * r[4] / mrtsq);
    cep[14] = (gbar - thbar * (r[4] * r[4]) / mrtsq);
    cep[15] = -(thbar * r[0] * r[5] / mrtsq);
    cep[16] = -(thbar * r[1] * r[5] / mrtsq);
    cep[17] = -(thbar * r[2] * r[5] / mrtsq);
    cep[18] = -(thbar * r[3] * r[5] / mrtsq);
    cep[19] = -(thbar * r[4] * r[5] / mrtsq);
    cep[20] = (gbar - thbar * (r[5] * r[5]) / mrtsq);
  }
  return ok;
}

// Lower-triangle index of (row, col) in the 21-word [D] storage.
__device__ __host__ constexpr int tri(int r, int c) {
  return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r;
}

// B-bar column terms of node j at one point (bmod.f:16 with
// alpha=(2+eps)/3, beta=(1-eps)/3).
struct BCol {
  double bx, by, bz;     // dN/dx, dN/dy, dN/dz
  double ux0, ux1;       // xx, yy row entries of the u dof (zz = yy entry)
  double vy0, vy1;       // xx, yy row entries of the v dof (zz = xx entry)
  double wz0, wz2;       // xx(=yy), zz row entries of the w dof
};

// ---------------------------------------------------------------------
// Recovery: one workgroup = E consecutive elements of one block, lanes are
// (element, Gauss point) in phases 1-3 and (element, dof) in phase 4.
// ---------------------------------------------------------------------
template <int NN, int G, bool BBAR>
__global__ void __launch_bounds__(256)
    k_recover(DevModel M, const WorkItem* __restrict__ work, int step,
              int iter, int nonlinear, const double* __restrict__ du,
              int want_energy) {
  constexpr int T = 3 * NN;
  constexpr int E = rec_elems(G);
  __shared__ double s_dn[3][G][NN];
  __shared__ double s_w[G];
  __shared__ double s_ce[E][T];
  __shared__ double s_du[E][T];
  __shared__ double s_gam[E][G][10];
  __shared__ double s_vol[BBAR ? E : 1][25];
  __shared__ double s_sig[E][G][6];

  const WorkItem wi = work[blockIdx.x];
  const BlockDesc bd = M.blocks[wi.blk];
  const GaussTable* gt = M.tables + bd.table;
  const int tid = threadIdx.x;
  const int cnt = wi.count;

  for (int k = tid; k < 3 * G * NN; k += 256) {
    int d = k / (G * NN), rem = k % (G * NN);
    s_dn[d][rem / NN][rem % NN] = gt->dn[d][rem / NN][rem % NN];
  }
  if (tid < G) s_w[tid] = gt->w[tid];
  for (int k = tid; k < cnt * NN; k += 256) {
    int e = k / NN, j = k % NN;
    int node = M.conn[bd.conn_off + (int64_t)(wi.first + e) * NN + j];
    s_ce[e][j] = M.coords[3 * node];
    s_ce[e][NN + j] = M.coords[3 * node + 1];
    s_ce[e][2 * NN + j] = M.coords[3 * node + 2];
    s_du[e][j] = du[3 * node];
    s_du[e][NN + j] = du[3 * node + 1];
    s_du[e][2 * NN + j] = du[3 * node + 2];
  }
  __syncthreads();

  const int e = tid / G, g = tid % G;
  const bool active = e < cnt;
  // phase 1: coordinate Jacobian, determinant and inverse (jacob1)
  if (active) {
    double j11 = 0, j12 = 0, j13 = 0, j21 = 0, j22 = 0, j23 = 0, j31 = 0,
           j32 = 0, j33 = 0;
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      double x = s_ce[e][j], y = s_ce[e][NN + j], z = s_ce[e][2 * NN + j];
      double a = s_dn[0][g][j], b = s_dn[1][g][j], c = s_dn[2][g][j];
      j11 = j11 + a * x; j12 = j12 + a * y; j13 = j13 + a * z;
      j21 = j21 + b * x; j22 = j22 + b * y; j23 = j23 + b * z;
      j31 = j31 + c * x; j32 = j32 + c * y; j33 = j33 + c * z;
    }
    double c1 = j22 * j33 - j23 * j32;
    double c2 = j21 * j33 - j23 * j31;
    double c3 = j21 * j32 - j22 * j31;
    double dj = j11 * c1 - j12 * c2 + j13 * c3;
    double* ga = s_gam[e][g];
    ga[0] = c1 / dj;
    ga[3] = -c2 / dj;
    ga[6] = c3 / dj;
    ga[1] = (j32 * j13 - j12 * j33) / dj;
    ga[4] = (j11 * j33 - j31 * j13) / dj;
    ga[7] = (j12 * j31 - j11 * j32) / dj;
    ga[2] = (j12 * j23 - j13 * j22) / dj;
    ga[5] = (j13 * j21 - j11 * j23) / dj;
    ga[8] = (j11 * j22 - j12 * j21) / dj;
    ga[9] = dj;
    if (dj <= 1.0e-20) atomicOr(&M.flags[2], 1);
  }
  __syncthreads();
  // phase 2: mean-dilatation volume terms (vol_terms + vol_avg)
  if (BBAR) {
    if (tid < cnt * 8) {
      int ee = tid >> 3, j = tid & 7;
      double vx = 0, vy = 0, vz = 0, volume = 0;
#pragma unroll
      for (int gg = 0; gg < G; ++gg) {
        const double* a = s_gam[ee][gg];
        double d = a[9];
        double nx = s_dn[0][gg][j], ny = s_dn[1][gg][j], nz = s_dn[2][gg][j];
        vx = vx + (a[0] * nx + a[1] * ny + a[2] * nz) * d;
        vy = vy + (a[3] * nx + a[4] * ny + a[5] * nz) * d;
        vz = vz + (a[6] * nx + a[7] * ny + a[8] * nz) * d;
        volume = volume + d;
      }
      double voli = 1.0 / volume;
      s_vol[ee][3 * j] = vx * voli;
      s_vol[ee][3 * j + 1] = vy * voli;
      s_vol[ee][3 * j + 2] = vz * voli;
    }
    __syncthreads();
  }
  const double third = 1.0 / 3.0;
  const double alpha = (2.0 + M.eps_bbar) * third;
  const double beta_b = (1.0 - M.eps_bbar) * third;

  // phase 3: strain increment, material update, state write-back
  if (active) {
    const double* ga = s_gam[e][g];
    double deps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      double nx = s_dn[0][g][j], ny = s_dn[1][g][j], nz = s_dn[2][g][j];
      double bx = ga[0] * nx + ga[1] * ny + ga[2] * nz;
      double by = ga[3] * nx + ga[4] * ny + ga[5] * nz;
      double bz = ga[6] * nx + ga[7] * ny + ga[8] * nz;
      double du_u = s_du[e][j], du_v = s_du[e][NN + j],
             du_w = s_du[e][2 * NN + j];
      if (BBAR) {
        double vx = s_vol[e][3 * j], vy = s_vol[e][3 * j + 1],
               vz = s_vol[e][3 * j + 2];
        double u0 = alpha * bx + beta_b * vx, u1 = beta_b * (vx - bx);
        double v0 = beta_b * (vy - by), v1 = alpha * by + beta_b * vy;
        double w0 = beta_b * (vz - bz), w2 = alpha * bz + beta_b * vz;
        deps[0] = deps[0] + u0 * du_u + v0 * du_v + w0 * du_w;
        deps[1] = deps[1] + u1 * du_u + v1 * du_v + w0 * du_w;
        deps[2] = deps[2] + u1 * du_u + v0 * du_v + w2 * du_w;
      } else {
        deps[0] = deps[0] + bx * du_u;
        deps[1] = deps[1] + by * du_v;
        deps[2] = deps[2] + bz * du_w;
      }
      deps[3] = deps[3] + by * du_u + bx * du_v;
      deps[4] = deps[4] + bz * du_v + by * du_w;
      deps[5] = deps[5] + bz * du_u + bx * du_w;
    }
    const int64_t SG = (int64_t)bd.span * G;
    const int64_t q = (int64_t)(wi.first + e) * G + g;
    double* sn = M.stress_n + 9 * bd.q_off + q;
    double* sn1 = M.stress_n1 + 9 * bd.q_off + q;
    double* en = M.strain_n + 6 * bd.q_off + q;
    double* en1 = M.strain_n1 + 6 * bd.q_off + q;
    double* cp = M.cep + 21 * bd.q_off + q;
#pragma unroll
    for (int k = 0; k < 6; ++k) en1[k * SG] = en[k * SG] + deps[k];
    double cgn[9], cgn1[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) cgn[k] = sn[k * SG];
    const MatDesc md = M.mats[bd.mat];
    double cep[21];
    if (nonlinear) {
      double h[kMM01Hist], h1[kMM01Hist];
      double* hn = M.hist_n + bd.hist_off + q;
      double* hn1 = M.hist_n1 + bd.hist_off + q;
      double f11 = 0.0;
      if (step != 1) {
#pragma unroll
        for (int k = 0; k < kMM01Hist; ++k) h[k] = hn[k * SG];
        f11 = M.hist_n[bd.hist_off + 10 * SG + g];  // history(1,11)
      }
      if (!mm01_point(step, md, cgn, cgn1, deps, h, h1, f11, cep))
        atomicOr(&M.flags[0], 1);
      if (iter > 0)  // rplstr scatters history only for iter > 0
#pragma unroll
        for (int k = 0; k < kMM01Hist; ++k) hn1[k * SG] = h1[k];
    } else {
      // drive_01_update_a / _b: linear-elastic [D]
      const double one = 1.0, two = 2.0, e_ = md.e, nu = md.nu;
      double c1 = (e_ / ((one + nu) * (one - two * nu)));
      double c2 = (one - nu) * c1;
      double c3 = ((one - two * nu) / two) * c1;
      double c4 = nu * c1;
      double dm[6][6] = {{c2, c4, c4, 0, 0, 0}, {c4, c2, c4, 0, 0, 0},
                         {c4, c4, c2, 0, 0, 0}, {0, 0, 0, c3, 0, 0},
                         {0, 0, 0, 0, c3, 0},   {0, 0, 0, 0, 0, c3}};
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        double s = cgn[k];
#pragma unroll
        for (int m = 0; m < 6; ++m) s = s + dm[m][k] * deps[m];
        cgn1[k] = s;
      }
      cgn1[6] = cgn1[7] = cgn1[8] = 0.0;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) cep[tri(r, c)] = dm[r][c];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) sn1[k * SG] = cgn1[k];
#pragma unroll
    for (int k = 0; k < 21; ++k) cp[k * SG] = cep[k];
#pragma unroll
    for (int k = 0; k < 6; ++k) s_sig[e][g][k] = cgn1[k];
    if (want_energy) {
      double factor = ga[9] * 1.0 * s_w[g];
      int64_t qq = bd.q_off + q;
      M.energy_q[2 * qq] = cgn1[6] * factor;
      M.energy_q[2 * qq + 1] = cgn1[7] * factor;
    }
  }
  __syncthreads();

  // phase 4: element internal forces f = sum_gp B^T sigma w |J| (gpifv1)
  const double scalar = bd.etype == W3D_TET10 ? 1.0 / 6.0 : 1.0;
  for (int k = tid; k < cnt * T; k += 256) {
    int ee = k / T, dof = k % T;
    int c = dof / NN, j = dof % NN;
    double f = 0.0;
    for (int gg = 0; gg < G; ++gg) {
      const double* ga = s_gam[ee][gg];
      double nx = s_dn[0][gg][j], ny = s_dn[1][gg][j], nz = s_dn[2][gg][j];
      double bx = ga[0] * nx + ga[1] * ny + ga[2] * nz;
      double by = ga[3] * nx + ga[4] * ny + ga[5] * nz;
      double bz = ga[6] * nx + ga[7] * ny + ga[8] * nz;
      const double* s = s_sig[ee][gg];
      double w = s_w[gg] * scalar, d = ga[9];
      if (c == 0) {
        double b0 = bx, b1 = 0.0;
        if (BBAR) {
          double vx = s_vol[ee][3 * j];
          b0 = alpha * bx + beta_b * vx;
          b1 = beta_b * (vx - bx);
        }
        f = f + b0 * s[0] * w * d + b1 * s[1] * w * d + b1 * s[2] * w * d +
            by * s[3] * w * d + bz * s[5] * w * d;
      } else if (c == 1) {
        double b1 = by, b0 = 0.0;
        if (BBAR) {
          double vy = s_vol[ee][3 * j + 1];
          b0 = beta_b * (vy - by);
          b1 = alpha * by + beta_b * vy;
        }
        f = f + b0 * s[0] * w * d + b1 * s[1] * w * d + b0 * s[2] * w * d +
            bx * s[3] * w * d + bz * s[4] * w * d;
      } else {
        double b2 = bz, b0 = 0.0;
        if (BBAR) {
          double vz = s_vol[ee][3 * j + 2];
          b0 = beta_b * (vz - bz);
          b2 = alpha * bz + beta_b * vz;
        }
        f = f + b0 * s[0] * w * d + b0 * s[1] * w * d + b2 * s[2] * w * d +
            by * s[4] * w * d + bx * s[5] * w * d;
      }
    }
    M.einfvec[bd.ifv_off + (int64_t)(wi.first + ee) * T + dof] = f;
  }
  for (int k = tid; k < cnt; k += 256) {
    double v = 0.0;
    for (int gg = 0; gg < G; ++gg) v = v + (s_w[gg] * scalar) * s_gam[k][gg][9];
    M.evol[(int64_t)bd.felem - 1 + wi.first + k] = v;
  }
}

// ---------------------------------------------------------------------
// Block tangent, vector FMA form.  Lanes own 4x4 tiles of the upper
// triangle of Ke; B and D*B of E elements at one Gauss point are staged
// in LDS per point; the per-point accumulation follows bdbtgen.
// ---------------------------------------------------------------------
template <int NN>
struct TanShape {
  static constexpr int T = 3 * NN;
  static constexpr int NT = (T + 3) / 4;
  static constexpr int L = NT * (NT + 1) / 2;
  static constexpr int E = tan_elems(NN);
};

template <int NN, int G, bool BBAR, typename KT>
__global__ void __launch_bounds__(256)
    k_tangent_vec(DevModel M, const WorkItem* __restrict__ work,
                  KT* __restrict__ ek_out) {
  using S = TanShape<NN>;
  constexpr int T = S::T, NT = S::NT, L = S::L, E = S::E, TP = 4 * NT;
  __shared__ double s_dn[3][G][NN];
  __shared__ double s_w[G];
  __shared__ double s_ce[E][T];
  __shared__ double s_gam[E][G][10];
  __shared__ double s_vol[BBAR ? E : 1][25];
  __shared__ double s_b[E][6][TP];
  __shared__ double s_db[E][6][TP];
  __shared__ double s_d[E][21];

  const WorkItem wi = work[blockIdx.x];
  const BlockDesc bd = M.blocks[wi.blk];
  const GaussTable* gt = M.tables + bd.table;
  const int tid = threadIdx.x;
  const int cnt = wi.count;
  const double scalar = bd.etype == W3D_TET10 ? 1.0 / 6.0 : 1.0;

  for (int k = tid; k < 3 * G * NN; k += 256) {
    int d = k / (G * NN), rem = k % (G * NN);
    s_dn[d][rem / NN][rem % NN] = gt->dn[d][rem / NN][rem % NN];
  }
  if (tid < G) s_w[tid] = gt->w[tid];
  for (int k = tid; k < cnt * NN; k += 256) {
    int e = k / NN, j = k % NN;
    int node = M.conn[bd.conn_off + (int64_t)(wi.first + e) * NN + j];
    s_ce[e][j] = M.coords[3 * node];
    s_ce[e][NN + j] = M.coords[3 * node + 1];
    s_ce[e][2 * NN + j] = M.coords[3 * node + 2];
  }
  __syncthreads();
  for (int k = tid; k < cnt * G; k += 256) {
    int e = k / G, g = k % G;
    double j11 = 0, j12 = 0, j13 = 0, j21 = 0, j22 = 0, j23 = 0, j31 = 0,
           j32 = 0, j33 = 0;
    for (int j = 0; j < NN; ++j) {
      double x = s_ce[e][j], y = s_ce[e][NN + j], z = s_ce[e][2 * NN + j];
      double a = s_dn[0][g][j], b = s_dn[1][g][j], c = s_dn[2][g][j];
      j11 = j11 + a * x; j12 = j12 + a * y; j13 = j13 + a * z;
      j21 = j21 + b * x; j22 = j22 + b * y; j23 = j23 + b * z;
      j31 = j31 + c * x; j32 = j32 + c * y; j33 = j33 + c * z;
    }
    double c1 = j22 * j33 - j23 * j32;
    double c2 = j21 * j33 - j23 * j31;
    double c3 = j21 * j32 - j22 * j31;
    double dj = j11 * c1 - j12 * c2 + j13 * c3;
    double* ga = s_gam[e][g];
    ga[0] = c1 / dj;
    ga[3] = -c2 / dj;
    ga[6] = c3 / dj;
    ga[1] = (j32 * j13 - j12 * j33) / dj;
    ga[4] = (j11 * j33 - j31 * j13) / dj;
    ga[7] = (j12 * j31 - j11 * j32) / dj;
    ga[2] = (j12 * j23 - j13 * j22) / dj;
    ga[5] = (j13 * j21 - j11 * j23) / dj;
    ga[8] = (j11 * j22 - j12 * j21) / dj;
    ga[9] = dj;
  }
  __syncthreads();
  const double third = 1.0 / 3.0;
  const double alpha = (2.0 + M.eps_bbar) * third;
  const double beta_b = (1.0 - M.eps_bbar) * third;
  if (BBAR) {
    if (tid < cnt * 8) {
      int ee = tid >> 3, j = tid & 7;
      double vx = 0, vy = 0, vz = 0, volume = 0;
      for (int gg = 0; gg < G; ++gg) {
        const double* a = s_gam[ee][gg];
        double d = a[9];
        double nx = s_dn[0][gg][j], ny = s_dn[1][gg][j], nz = s_dn[2][gg][j];
        vx = vx + (a[0] * nx + a[1] * ny + a[2] * nz) * d;
        vy = vy + (a[3] * nx + a[4] * ny + a[5] * nz) * d;
        vz = vz + (a[6] * nx + a[7] * ny + a[8] * nz) * d;
        volume = volume + d;
      }
      double voli = 1.0 / volume;
      s_vol[ee][3 * j] = vx * voli;
      s_vol[ee][3 * j + 1] = vy * voli;
      s_vol[ee][3 * j + 2] = vz * voli;
    }
    __syncthreads();
  }

  // tile of this lane
  const int e = tid / L, t = tid % L;
  int I = 0, J = 0;
  {
    int rem = t;
    // upper tiles enumerated column by column: (0,0),(0,1),(1,1),(0,2),...
    while (rem > J) {
      rem -= J + 1;
      ++J;
    }
    I = rem;
  }
  const bool owner = e < cnt;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[a][c] = 0.0;

  const int64_t SG = (int64_t)bd.span * G;
  for (int g = 0; g < G; ++g) {
    // [D] of every element at this point, scaled by w |J| (and 1/6)
    for (int k = tid; k < cnt * 21; k += 256) {
      int ee = k / 21, c = k % 21;
      int64_t q = (int64_t)(wi.first + ee) * G + g;
      double f = s_w[g] * s_gam[ee][g][9];
      double v = f * M.cep[21 * bd.q_off + c * SG + q];
      if (bd.etype == W3D_TET10) v = scalar * v;
      s_d[ee][c] = v;
    }
    // B columns (zero padded to TP)
    for (int k = tid; k < cnt * TP; k += 256) {
      int ee = k / TP, dof = k % TP;
      double col[6] = {0, 0, 0, 0, 0, 0};
      if (dof < T) {
        int c = dof / NN, j = dof % NN;
        const double* ga = s_gam[ee][g];
        double nx = s_dn[0][g][j], ny = s_dn[1][g][j], nz = s_dn[2][g][j];
        double bx = ga[0] * nx + ga[1] * ny + ga[2] * nz;
        double by = ga[3] * nx + ga[4] * ny + ga[5] * nz;
        double bz = ga[6] * nx + ga[7] * ny + ga[8] * nz;
        if (c == 0) {
          col[0] = bx; col[3] = by; col[5] = bz;
          if (BBAR) {
            double vx = s_vol[ee][3 * j];
            col[0] = alpha * bx + beta_b * vx;
            col[1] = col[2] = beta_b * (vx - bx);
          }
        } else if (c == 1) {
          col[1] = by; col[3] = bx; col[4] = bz;
          if (BBAR) {
            double vy = s_vol[ee][3 * j + 1];
            col[0] = col[2] = beta_b * (vy - by);
            col[1] = alpha * by + beta_b * vy;
          }
        } else {
          col[2] = bz; col[4] = by; col[5] = bx;
          if (BBAR) {
            double vz = s_vol[ee][3 * j + 2];
            col[0] = col[1] = beta_b * (vz - bz);
            col[2] = alpha * bz + beta_b * vz;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 6; ++r) s_b[ee][r][dof] = col[r];
    }
    __syncthreads();
    for (int k = tid; k < cnt * TP; k += 256) {
      int ee = k / TP, dof = k % TP;
      double b[6];
#pragma unroll
      for (int r = 0; r < 6; ++r) b[r] = s_b[ee][r][dof];
      const double* d = s_d[ee];
#pragma unroll
      for (int kk = 0; kk < 6; ++kk)
        s_db[ee][kk][dof] = d[tri(0, kk)] * b[0] + d[tri(1, kk)] * b[1] +
                            d[tri(2, kk)] * b[2] + d[tri(3, kk)] * b[3] +
                            d[tri(4, kk)] * b[4] + d[tri(5, kk)] * b[5];
    }
    __syncthreads();
    if (owner) {
#pragma unroll
      for (int r = 0; r < 6; ++r) {
        double bc[4], dr[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) bc[c] = s_b[e][r][4 * J + c];
#pragma unroll
        for (int a = 0; a < 4; ++a) dr[a] = s_db[e][r][4 * I + a];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[a][c] = acc[a][c] + bc[c] * dr[a];
      }
    }
    __syncthreads();
  }
  if (owner) {
    KT* out = ek_out + bd.ek_off + (int64_t)(wi.first + e) * bd.utsz;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        int row = 4 * I + a, col = 4 * J + c;
        if (row <= col && col < T) out[col * (col + 1) / 2 + row] = (KT)acc[a][c];
      }
  }
}

// ---------------------------------------------------------------------
// Gathers (deterministic: each output sums its inputs in a fixed order)
// ---------------------------------------------------------------------
__global__ void k_ifv_gather(const double* __restrict__ einfvec,
                             const int64_t* __restrict__ inv_ptr,
                             const int64_t* __restrict__ inv_off,
                             const int32_t* __restrict__ inv_nn, int nonode,
                             double* __restrict__ ifv) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nonode) return;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int64_t p = inv_ptr[n]; p < inv_ptr[n + 1]; ++p) {
    int64_t o = inv_off[p];
    int nn = inv_nn[p];
    fx = fx + einfvec[o];
    fy = fy + einfvec[o + nn];
    fz = fz + einfvec[o + 2 * nn];
  }
  ifv[3 * n] = fx;
  ifv[3 * n + 1] = fy;
  ifv[3 * n + 2] = fz;
}

// Per-block energy and plastic work in rstgp1_b order (point outer,
// element inner); one thread per block.
__global__ void k_energy_reduce(const double* __restrict__ energy_q,
                                const int64_t* __restrict__ blk_q_off,
                                const int32_t* __restrict__ blk_ngp,
                                const int32_t* __restrict__ blk_span,
                                int nblk, double* __restrict__ sums) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  double en = 0.0, pw = 0.0;
  int G = blk_ngp[b], S = blk_span[b];
  int64_t q0 = blk_q_off[b];
  for (int g = 0; g < G; ++g)
    for (int i = 0; i < S; ++i) {
      int64_t q = q0 + (int64_t)i * G + g;
      en = en + energy_q[2 * q];
      pw = pw + energy_q[2 * q + 1];
    }
  sums[2 * b] = en;
  sums[2 * b + 1] = pw;
}

// Row-by-row assembly (assem_a_row): output t sums its Ke words in the
// order of the elements on the row's node.
template <typename KT>
__global__ void k_assemble(const KT* __restrict__ ek,
                           const int64_t* __restrict__ row_ptr,
                           const uint32_t* __restrict__ src, int64_t nout,
                           double* __restrict__ out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout) return;
  double s = 0.0;
  for (int64_t p = row_ptr[t]; p < row_ptr[t + 1]; ++p)
    s = s + (double)ek[src[p]];
  out[t] = s;
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
int group_key(int nnode, int ngp, int bbar) {
  return (nnode * 100 + ngp) * 2 + (bbar ? 1 : 0);
}

#define W3D_GROUPS(X)    \
  X(8, 8, true)          \
  X(8, 8, false)         \
  X(20, 27, false)       \
  X(20, 8, false)        \
  X(20, 14, false)       \
  X(20, 9, false)        \
  X(10, 4, false)        \
  X(10, 1, false)        \
  X(10, 5, false)

int launch_recover(const DevModel& M, const std::vector<BlockDesc>& hblocks,
                   const WorkItem* d_work, const std::vector<int>& groups,
                   const std::vector<int>& gstart,
                   const std::vector<int>& gcount, int step, int iter,
                   int nonlinear, const double* d_du, int want_energy,
                   hipStream_t s) {
  (void)hblocks;
  for (size_t k = 0; k < groups.size(); ++k) {
    if (gcount[k] == 0) continue;
    const WorkItem* w = d_work + gstart[k];
    bool done = false;
#define X(NN, G, BB)                                                       \
  if (!done && groups[k] == group_key(NN, G, BB)) {                         \
    hipLaunchKernelGGL((k_recover<NN, G, BB>), dim3(gcount[k]), dim3(256), \
                       0, s, M, w, step, iter, nonlinear, d_du,            \
                       want_energy);                                       \
    done = true;                                                           \
  }
    W3D_GROUPS(X)
#undef X
    if (!done) return W3D_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_tangent(const DevModel& M, const std::vector<BlockDesc>& hblocks,
                   const WorkItem* d_work, const std::vector<int>& groups,
                   const std::vector<int>& gstart,
                   const std::vector<int>& gcount, int precision,
                   hipStream_t s) {
  (void)hblocks;
  for (size_t k = 0; k < groups.size(); ++k) {
    if (gcount[k] == 0) continue;
    const WorkItem* w = d_work + gstart[k];
    bool done = false;
#define X(NN, G, BB)                                                         \
  if (!done && groups[k] == group_key(NN, G, BB)) {                           \
    if (precision == W3D_FP32)                                               \
      hipLaunchKernelGGL((k_tangent_vec<NN, G, BB, float>), dim3(gcount[k]), \
                         dim3(256), 0, s, M, w, M.ek32);                     \
    else                                                                     \
      hipLaunchKernelGGL((k_tangent_vec<NN, G, BB, double>),                 \
                         dim3(gcount[k]), dim3(256), 0, s, M, w, M.ek64);    \
    done = true;                                                             \
  }
    W3D_GROUPS(X)
#undef X
    if (!done) return W3D_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_ifv_gather(const double* einfvec, const int64_t* inv_ptr,
                      const int64_t* inv_off, const int32_t* inv_nn,
                      int nonode, double* ifv, hipStream_t s) {
  int nb = (nonode + 255) / 256;
  if (nb > 0)
    hipLaunchKernelGGL(k_ifv_gather, dim3(nb), dim3(256), 0, s, einfvec,
                       inv_ptr, inv_off, inv_nn, nonode, ifv);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_energy_reduce(const double* energy_q, const int64_t* blk_q_off,
                         const int32_t* blk_ngp, const int32_t* blk_span,
                         int nblk, double* sums, hipStream_t s) {
  int nb = (nblk + 63) / 64;
  if (nb > 0)
    hipLaunchKernelGGL(k_energy_reduce, dim3(nb), dim3(64), 0, s, energy_q,
                       blk_q_off, blk_ngp, blk_span, nblk, sums);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_assemble(const double* ek64, const float* ek32, int fp32,
                    const int64_t* row_ptr, const uint32_t* src, int64_t nout,
                    double* out, hipStream_t s) {
  int64_t nb = (nout + 255) / 256;
  if (nb > 0) {
    if (fp32)
      hipLaunchKernelGGL(k_assemble<float>, dim3((unsigned)nb), dim3(256), 0,
                         s, ek32, row_ptr, src, nout, out);
    else
      hipLaunchKernelGGL(k_assemble<double>, dim3((unsigned)nb), dim3(256), 0,
                         s, ek64, row_ptr, src, nout, out);
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

}  // namespace w3d
