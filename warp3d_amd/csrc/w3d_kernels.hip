// w3d_kernels.hip -- gfx950 kernels of the element-block hot path:
// strain/stress recovery + mm01 + element internal forces, and the
// deterministic gathers that assemble element forces and block tangents.
//
// Reference routines restated (file:line under the reference's src/):
//   rknstr.f:563 rknstr_sm_displ, gtmat1.f:300 jacob1, vol_terms.f:14,
//   vol_avg.f:13, blcmp1.f:15, bmod.f:16, gtlsn1.f:25, rstgp1.f:351 rstgp2,
//   rstgp1.f:614 drive_01_update (+ :998 drive_01_update_b), mm01.f:24-1348
//   (mm01_set_history, mm01_init, mm01_general, mm01_sig_final, mm01_wrapup,
//   cnst1), rstgp1.f:5360 rstgp1_store_cep, rstgp1.f:5189 rstgp1_b,
//   gpifv1.f:16, rknifv.f:141, addifv.f:16, assemble_code.f:659
//   assem_by_row.
#include <hip/hip_runtime.h>

#include "w3d_geonl.h"
#include "w3d_internal.h"

namespace w3d {

// ---------------------------------------------------------------------
// Material properties of one point: the block's (setup_mm01_rknstr,
// rknstr.f:1018: REAL props widened, H' formed in REAL) unless the
// material has a segmental curve set: set_up_segmental
// (setup_mat_props.f:67) interpolates e, nu and the isotropic alpha at the
// point's temperatures at n+1 and n and the curve stresses at n+1,
// set_up_h_prime (:406) takes H' and the yield stress from the curve's
// first two points.
// ---------------------------------------------------------------------
struct PointProps {
  double e, nu, e_n, nu_n, hprime, yld;
  double alpha[6], alpha_n[6];
};

// linear_interpolate (setup_mat_props.f:361): constant outside the table
__device__ inline double lin_interp(double x, int n, const double* xs,
                                    const double* ys, int ystride) {
  if (x <= xs[0]) return ys[0];
  if (x >= xs[n - 1]) return ys[(n - 1) * ystride];
  int p = 1;
  while (p < n - 1 && x > xs[p]) ++p;
  const double x1 = xs[p - 1], x2 = xs[p], y1 = ys[(p - 1) * ystride],
               y2 = ys[p * ystride];
  return y1 + (x - x1) * (y2 - y1) / (x2 - x1);
}

__device__ inline void point_props(const MatDesc& md, const ThermDesc* td,
                                   double t_n1, double dtemp,
                                   PointProps& p) {
  p.e = p.e_n = md.e;
  p.nu = p.nu_n = md.nu;
  p.hprime = md.hprime;
  p.yld = md.yld;
#pragma unroll
  for (int k = 0; k < 6; ++k) p.alpha[k] = p.alpha_n[k] = td->alpha[k];
  if (td->curve_type == 0) return;
  double s1 = td->sig[0][0], s2 = td->sig[0][1];
  if (td->curve_type == 2) {
    const int n = td->ncurves;
    const double t_n = t_n1 - dtemp;
    p.e = lin_interp(t_n1, n, td->temp, td->e, 1);
    p.e_n = lin_interp(t_n, n, td->temp, td->e, 1);
    p.nu = lin_interp(t_n1, n, td->temp, td->nu, 1);
    p.nu_n = lin_interp(t_n, n, td->temp, td->nu, 1);
    const double a1 = lin_interp(t_n1, n, td->temp, td->alpha_c, 1);
    const double a0 = lin_interp(t_n, n, td->temp, td->alpha_c, 1);
    p.alpha[0] = p.alpha[1] = p.alpha[2] = a1;
    p.alpha_n[0] = p.alpha_n[1] = p.alpha_n[2] = a0;
    s1 = lin_interp(t_n1, n, td->temp, &td->sig[0][0], 2);
    s2 = lin_interp(t_n1, n, td->temp, &td->sig[0][1], 2);
  }
  p.hprime = (s2 - s1) / (td->curve_eps[1] - td->curve_eps[0]);
  p.yld = s1;
}

// gauss_pt_temps (gp_temp_eps.f:17) of lane (element le of the item, point
// g): the step's temperature change (zero unless the step imposes one) and
// the temperature at n+1, from the element's node and element values
// (gadtemps / gatemps, gatemps.f:63,117)
template <int NN>
__device__ inline void point_temps(const DevModel& M, const BlockDesc& bd,
                                   const GaussTable* gt, int elem_in_blk,
                                   int g, double& t_n1, double& dtemp) {
  const int elem = bd.felem - 1 + elem_in_blk;  // 0-based
  const double* tn = M.temps;
  const double* dtn = M.temps + M.nonode;
  const double te = M.temps[2 * M.nonode + elem];
  const double dte = M.temps[2 * M.nonode + M.noelem + elem];
  const int32_t* cn = M.conn + bd.conn_off + (int64_t)elem_in_blk * NN;
  double a = 0.0, b = 0.0;
#pragma unroll 4
  for (int j = 0; j < NN; ++j) {
    const int node = cn[j];
    const double sf = gt->sf[g][j];
    const double d = M.temps_incr ? dtn[node] + dte : 0.0;
    a = a + sf * d;
    b = b + sf * (tn[node] + te + d);
  }
  dtemp = M.temps_incr ? a : 0.0;
  t_n1 = b;
}

// gp_temp_eps (gp_temp_eps.f:163): adds the (negative) thermal strain
// increment to the displacement strain increment deps
__device__ inline void thermal_strain(const double* alpha_n1,
                                      const double* alpha_n,
                                      double alpha_zero, double t_n1,
                                      double dtemp, double* deps) {
  const double tn = (t_n1 - dtemp) - alpha_zero, tn1 = t_n1 - alpha_zero;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double th = alpha_n1[k] * tn1 - alpha_n[k] * tn;
    deps[k] = deps[k] + (0.0 - th);
  }
}

// ---------------------------------------------------------------------
// mm01 (bilinear Mises) + cnst1 for one Gauss point.
// cgn: 9 stress words at n (words 8, 9 may be reset at step 1); h: 11
// history words at n; pt: the point's moduli at n+1 and n, H' and yield
// stress.  The results go straight to the point's slots (stride SG):
// stresses to sn1, history to hn1 (when store_hist), the 21-word [D] to
// cp; sig receives stress words 1-8 for the caller.  Returns false on a
// negative plastic multiplier (die_abort in mm01_general).
// ---------------------------------------------------------------------
__device__ inline double state_word(int st) {
  return __longlong_as_double((long long)(unsigned)st);
}

__device__ bool mm01_point(int step, const MatDesc& md, const PointProps& pt,
                           double* cgn, const double* deps, double* h,
                           double first_hist11, double* __restrict__ sn1,
                           double* __restrict__ hn1, double* __restrict__ cp,
                           int64_t SG, bool store_hist, double* sig) {
  const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, half = 0.5,
               six = 6.0, four = 4.0;
  const double e = pt.e, nu = pt.nu, beta = md.beta, hprime = pt.hprime,
               yld = pt.yld;
  const double e_n = pt.e_n, nu_n = pt.nu_n;
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
  double g_n = e_n / two / (one + nu_n);
  double een1 = (cgn[0] - nu_n * (cgn[1] + cgn[2])) / e_n;
  double een2 = (cgn[1] - nu_n * (cgn[0] + cgn[2])) / e_n;
  double een3 = (cgn[2] - nu_n * (cgn[0] + cgn[1])) / e_n;
  double eps_vol_n1 = een1 + een2 + een3 + deps_vol;
  double eps_mean_n = (een1 + een2 + een3) / three;
  double shear_mod = e / (two * (one + nu));
  double r[6];
  r[0] = two * shear_mod * ((een1 - eps_mean_n) + (deps[0] - eps_mean));
  r[1] = two * shear_mod * ((een2 - eps_mean_n) + (deps[1] - eps_mean));
  r[2] = two * shear_mod * ((een3 - eps_mean_n) + (deps[2] - eps_mean));
  r[3] = shear_mod * (cgn[3] / g_n + deps[3]);
  r[4] = shear_mod * (cgn[4] / g_n + deps[4]);
  r[5] = shear_mod * (cgn[5] / g_n + deps[5]);
  // r holds the elastic trial deviator here
  double hbari_np1 = beta * hprime;
  double hbark_np1 = (one - beta) * hprime;
  double hbark_n = (one - beta) * h[4];
  double kbar = (yld + hbari_np1 * h[2]) / root3;
  double lk = one;
  if (fabs(hbark_n) > htol) lk = hbark_np1 / hbark_n;
  double t0 = r[0] - h[5] * lk, t1 = r[1] - h[6] * lk, t2 = r[2] - h[7] * lk,
         t3 = r[3] - h[8] * lk, t4 = r[4] - h[9] * lk, t5 = r[5] - h[10] * lk;
  double yf = sqrt(t0 * t0 + t1 * t1 + t2 * t2 +
                   two * (t3 * t3 + t4 * t4 + t5 * t5)) - root2 * kbar;
#pragma unroll
  for (int k = 0; k < 6; ++k) r[k] = r[k] - h[5 + k];
  double mrts = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2] +
                     two * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]));
  const bool yield = !md.lnelas && yf >= yld_tol * root2 * kbar;
  double dev[6];
  double ldt = zero, k_n1, ebarp_n1;
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
    double q1 = a * a * vbar - two * d * d;
    double q2 = -four * d * kbar - two * a * vbar - two * a * b * vwbar;
    double q3 = b * b * wbar + two * b * vwbar + vbar - two * kbar * kbar;
    double discr = q2 * q2 - four * q1 * q3;
    if (discr >= zero) {
      double sq = sqrt(discr);
      double x1 = (-q2 + sq) / (two * q1);
      double x2 = (-q2 - sq) / (two * q1);
      ldt = x1 < x2 ? x1 : x2;
    }
    if (ldt < zero) ok = false;
    k_n1 = kbar + (r2 / three) * hbari_np1 * ldt;
    ebarp_n1 = h[2] + ldt * root23;
    double const1 = twthrd * hbark_np1 * ldt / mrts;
    double const2 = one - ldt * (twthrd * hbark_np1 + two * shear_mod) / mrts;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double alpha_n1 = lk * h[5 + k] + const1 * r[k];
      if (store_hist) hn1[(5 + k) * SG] = alpha_n1;
      dev[k] = alpha_n1 + b * h[5 + k] + const2 * r[k];
    }
  } else {
    k_n1 = md.lnelas ? zero : kbar;
    ebarp_n1 = h[2];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      if (store_hist) hn1[(5 + k) * SG] = h[5 + k] * lk;
      dev[k] = r[k] + h[5 + k];
    }
  }
  if (store_hist) {
    hn1[0] = ldt;
    hn1[SG] = k_n1;
    hn1[2 * SG] = ebarp_n1;
    hn1[3 * SG] = state_word(yield ? 1 : 3);
    hn1[4 * SG] = hprime;
  }
  // mm01_sig_final
  double sig_mean =
      eps_vol_n1 *
      (three * e * nu / ((one + nu) * (one - two * nu)) + two * shear_mod) /
      three;
  double c1[6];
  c1[0] = dev[0] + sig_mean;
  c1[1] = dev[1] + sig_mean;
  c1[2] = dev[2] + sig_mean;
  c1[3] = dev[3];
  c1[4] = dev[4];
  c1[5] = dev[5];
  double en1 = cgn[6] + half * (deps[0] * (c1[0] + cgn[0]) +
                                deps[1] * (c1[1] + cgn[1]) +
                                deps[2] * (c1[2] + cgn[2]) +
                                deps[3] * (c1[3] + cgn[3]) +
                                deps[4] * (c1[4] + cgn[4]) +
                                deps[5] * (c1[5] + cgn[5]));
  // mm01_wrapup
  double pw1 = cgn[7], ep1 = cgn[8];
  if (yield) {
    const double* s = cgn;
    double sbn = sqrt((s[0] - s[1]) * (s[0] - s[1]) +
                      (s[1] - s[2]) * (s[1] - s[2]) +
                      (s[0] - s[2]) * (s[0] - s[2]) +
                      six * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])) /
                 root2;
    s = c1;
    double sbn1 = sqrt((s[0] - s[1]) * (s[0] - s[1]) +
                       (s[1] - s[2]) * (s[1] - s[2]) +
                       (s[0] - s[2]) * (s[0] - s[2]) +
                       six * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5])) /
                  root2;
    double depbar = ebarp_n1 - h[2];
    pw1 = cgn[7] + half * (sbn1 + sbn) * depbar;
    ep1 = cgn[8] + depbar;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    sn1[k * SG] = c1[k];
    sig[k] = c1[k];
  }
  sn1[6 * SG] = en1;
  sn1[7 * SG] = pw1;
  sn1[8 * SG] = ep1;
  sig[6] = en1;
  sig[7] = pw1;
  // cnst1, written straight into the stored lower triangle
  if (!yield) {
    double e1 = (e / ((one + nu) * (one - two * nu)));
    double c2 = (one - nu) * e1;
    double c3 = ((one - two * nu) / two) * e1;
    double c4 = nu * e1;
    const double v[21] = {c2,   c4,   c2,   c4,   c4,   c2,   zero,
                          zero, zero, c3,   zero, zero, zero, zero,
                          c3,   zero, zero, zero, zero, zero, c3};
#pragma unroll
    for (int k = 0; k < 21; ++k) cp[k * SG] = v[k];
  } else {
    const double root2c = 1.414213562;
    double g = e / (two * (one + nu));
    double l = (e * nu) / ((one + nu) * (one - two * nu));
    double kk = (three * l + two * g) / three;
    double mrtsq = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] +
                   two * (r[3] * r[3] + r[4] * r[4] + r[5] * r[5]);
    double bb = (root2c * k_n1 + (two / three) * (one - beta) * hprime * ldt) /
                sqrt(mrtsq);
    double gamma = one / (one + hprime / (three * g));
    double gambar = gamma - one + bb;
    double gbar = g * bb;
    double albar = kk - two * gbar / three;
    double thbar = two * g * gambar;
    cp[0] = (albar + two * gbar - thbar * (r[0] * r[0]) / mrtsq);
    cp[1 * SG] = (albar - thbar * r[0] * r[1] / mrtsq);
    cp[2 * SG] = (albar + two * gbar - thbar * (r[1] * r[1]) / mrtsq);
    cp[3 * SG] = (albar - thbar * r[0] * r[2] / mrtsq);
    cp[4 * SG] = (albar - thbar * r[2] * r[1] / mrtsq);
    cp[5 * SG] = (albar + two * gbar - thbar * (r[2] * r[2]) / mrtsq);
    cp[6 * SG] = -(thbar * r[0] * r[3] / mrtsq);
    cp[7 * SG] = -(thbar * r[1] * r[3] / mrtsq);
    cp[8 * SG] = -(thbar * r[2] * r[3] / mrtsq);
    cp[9 * SG] = (gbar - thbar * (r[3] * r[3]) / mrtsq);
    cp[10 * SG] = -(thbar * r[0] * r[4] / mrtsq);
    cp[11 * SG] = -(thbar * r[1] * r[4] / mrtsq);
    cp[12 * SG] = -(thbar * r[2] * r[4] / mrtsq);
    cp[13 * SG] = -(thbar * r[3] * r[4] / mrtsq);
    cp[14 * SG] = (gbar - thbar * (r[4] * r[4]) / mrtsq);
    cp[15 * SG] = -(thbar * r[0] * r[5] / mrtsq);
    cp[16 * SG] = -(thbar * r[1] * r[5] / mrtsq);
    cp[17 * SG] = -(thbar * r[2] * r[5] / mrtsq);
    cp[18 * SG] = -(thbar * r[3] * r[5] / mrtsq);
    cp[19 * SG] = -(thbar * r[4] * r[5] / mrtsq);
    cp[20 * SG] = (gbar - thbar * (r[5] * r[5]) / mrtsq);
  }
  return ok;
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
// MODE 0: mm01 blocks, everything fused.  MODE 1: mm10 blocks, phases 1-3
// up to the strain increment (left in uddt for the mm10 kernel).  MODE 2:
// mm10 blocks, energy and element forces from the updated stresses.
// GEONL (finite strains; rstgp1.f:19, gtmat1.f:17, rknifv.f:141): the
// strain increment is formed with the mid-step configuration X + u + du/2
// and rotated to the unrotated frame by [R n+1/2]; [R n+1] and det F n+1
// come from F of the n = 0 configuration; the forces use the n+1
// configuration and the Cauchy stress [R n+1] t [R n+1]^T.  With mm10
// (MODE 1) [R n+1] and the element's last n = 0 Jacobian are left for the
// mm10 and grad Fe kernels; MODE 2 forms [R n+1] again from the same
// inputs.
// ---------------------------------------------------------------------
template <int NN, int G, bool BBAR, int MODE, bool GEONL = false>
__global__ void __launch_bounds__(256)
    k_recover(DevModel M, const WorkItem* __restrict__ work, int step,
              int iter, int nonlinear, const double* __restrict__ u,
              const double* __restrict__ du, int want_energy,
              double* __restrict__ en_wg, int slot0) {
  constexpr int T = 3 * NN;
  constexpr int E = rec_elems(G);
  // coordinates, displacements and increments (phases 0-3) share their
  // LDS with the updated stresses and energy terms (phases 3-4)
  constexpr int NA = (GEONL ? 3 : 2) * E * T, NB = 8 * E * G;
  __shared__ double s_dn[3][G][NN];
  __shared__ double s_w[G];
  __shared__ double s_buf[NA > NB ? NA : NB];
  __shared__ double s_gam[E][G][10];
  __shared__ double s_vol[BBAR ? E : 1][25];
  __shared__ double s_red[4];
  double(*s_ce)[T] = reinterpret_cast<double(*)[T]>(s_buf);
  double(*s_du)[T] = reinterpret_cast<double(*)[T]>(s_buf + E * T);
  double(*s_u)[T] = reinterpret_cast<double(*)[T]>(s_buf + 2 * E * T);
  double(*s_sig)[G][6] = reinterpret_cast<double(*)[G][6]>(s_buf);
  double* s_en0 = s_buf + 6 * E * G;
  double* s_en1 = s_buf + 7 * E * G;

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
    // (the mm10 force pass needs the increments only for F n+1)
    if (MODE != 2 || GEONL) {
      s_du[e][j] = du[3 * node];
      s_du[e][NN + j] = du[3 * node + 1];
      s_du[e][2 * NN + j] = du[3 * node + 2];
    }
    if (GEONL) {
      s_u[e][j] = u[3 * node];
      s_u[e][NN + j] = u[3 * node + 1];
      s_u[e][2 * NN + j] = u[3 * node + 2];
    }
    // dupstr: displacements of a skewed node to global coordinates
    // (trnvec, switch 2)
    const int ti = bd.trn ? M.trn_idx[node] : -1;
    if (ti >= 0) {
      const double* t = M.trn_mat + 9 * ti;
      double v[3];
      if (MODE != 2 || GEONL) {
        v[0] = s_du[e][j]; v[1] = s_du[e][NN + j]; v[2] = s_du[e][2 * NN + j];
        trn_apply_t(t, v);
        s_du[e][j] = v[0]; s_du[e][NN + j] = v[1]; s_du[e][2 * NN + j] = v[2];
      }
      if (GEONL) {
        v[0] = s_u[e][j]; v[1] = s_u[e][NN + j]; v[2] = s_u[e][2 * NN + j];
        trn_apply_t(t, v);
        s_u[e][j] = v[0]; s_u[e][NN + j] = v[1]; s_u[e][2 * NN + j] = v[2];
      }
    }
  }
  __syncthreads();

  const int e = tid / G, g = tid % G;
  const bool active = e < cnt;
  // jacob1 of lane (e, g) on the configuration X + a u + b du; jout (when
  // given) receives the Jacobian itself, row-major
  auto jacobian = [&](double a_u, double b_du, double* jout) {
    double j11 = 0, j12 = 0, j13 = 0, j21 = 0, j22 = 0, j23 = 0, j31 = 0,
           j32 = 0, j33 = 0;
#pragma unroll 2
    for (int j = 0; j < NN; ++j) {
      double x = s_ce[e][j], y = s_ce[e][NN + j], z = s_ce[e][2 * NN + j];
      if (GEONL && a_u != 0.0) {
        x = x + s_u[e][j] + b_du * s_du[e][j];
        y = y + s_u[e][NN + j] + b_du * s_du[e][NN + j];
        z = z + s_u[e][2 * NN + j] + b_du * s_du[e][2 * NN + j];
      }
      double a = s_dn[0][g][j], b = s_dn[1][g][j], c = s_dn[2][g][j];
      j11 = j11 + a * x; j12 = j12 + a * y; j13 = j13 + a * z;
      j21 = j21 + b * x; j22 = j22 + b * y; j23 = j23 + b * z;
      j31 = j31 + c * x; j32 = j32 + c * y; j33 = j33 + c * z;
    }
    if (jout) {
      jout[0] = j11; jout[1] = j12; jout[2] = j13;
      jout[3] = j21; jout[4] = j22; jout[5] = j23;
      jout[6] = j31; jout[7] = j32; jout[8] = j33;
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
  };
  // finite strains: [R n+1], [R n+1/2], det F n+1 and det J at n = 0 of
  // the lane's point (gtmat1: F from the n = 0 Jacobian)
  double rn1[9], rnh[9], dfn1 = 0.0, dj0 = 0.0;
  if (GEONL && active) {
    double j0[9];
    jacobian(0.0, 0.0, j0);
    // mm10 finite strains: the Jacobian left in local_work%jac after the
    // element's last point, for the grad Fe fit (mm10_calc_grads)
    if (MODE == 1 && g == G - 1) {
      double* jt = M.jac_tmp + 9 * ((int64_t)bd.felem - 1 + wi.first + e);
#pragma unroll
      for (int k = 0; k < 9; ++k) jt[k] = j0[k];
    }
    const double* ga = s_gam[e][g];
    dj0 = ga[9];
    double g0[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) g0[k] = ga[k];
    double f[9];
    dfn1 = geonl_f<NN>(g0, s_dn[0][g], s_dn[1][g], s_dn[2][g], s_u[e],
                       s_du[e], 1.0, f);
    bool bad = dfn1 <= 1.0e-20;
    geonl_rotation(f, rn1);
    double dfh = geonl_f<NN>(g0, s_dn[0][g], s_dn[1][g], s_dn[2][g], s_u[e],
                             s_du[e], 0.5, f);
    bad = bad || dfh <= 1.0e-20;
    geonl_rotation(f, rnh);
    if (bad) atomicOr(&M.flags[3], 1);
    if (MODE == 1) {  // [R n+1] of the point for the mm10 kernel
      const int64_t SG = (int64_t)bd.span * G;
      const int64_t q = (int64_t)(wi.first + e) * G + g;
      double* rr = M.rot_tmp + 9 * bd.q_off + q;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) rr[(a + 3 * c) * SG] = rn1[3 * a + c];
    }
  }
  if (GEONL) __syncthreads();  // the n = 0 Jacobians are replaced below
  // phase 1: coordinate Jacobian, determinant and inverse (jacob1), on the
  // mid-step configuration for finite strains (rknstr_geonl)
  if (active) jacobian(GEONL ? 1.0 : 0.0, 0.5, nullptr);
  __syncthreads();
  // phase 2: mean-dilatation volume terms (vol_terms + vol_avg)
  auto vol_terms = [&]() {
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
  };
  if (BBAR) {
    vol_terms();
    __syncthreads();
  }
  const double third = 1.0 / 3.0;
  const double alpha = (2.0 + M.eps_bbar) * third;
  const double beta_b = (1.0 - M.eps_bbar) * third;

  // phase 3: strain increment, material update, state write-back
  double sig[8];
  if (MODE == 2 && active) {  // stresses left by the mm10 update
    const int64_t SG = (int64_t)bd.span * G;
    const int64_t q = (int64_t)(wi.first + e) * G + g;
    const double* sn1 = M.stress_n1 + 9 * bd.q_off + q;
#pragma unroll
    for (int k = 0; k < 8; ++k) sig[k] = sn1[k * SG];
  }
  if (MODE != 2 && active) {
    const double* ga = s_gam[e][g];
    double deps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll 2
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
    if (GEONL) {  // uddt = [qnhalf] ddt (qmply1)
      double qh[36], ddt[6];
      geonl_q(rnh, 1, qh);
#pragma unroll
      for (int k = 0; k < 6; ++k) ddt[k] = deps[k];
      geonl_qmul(qh, ddt, deps);
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
    // the point's temperatures and properties; the thermal strain
    // increment comes off the increment the material sees (the stored
    // strains above are the displacements')
    const MatDesc md = M.mats[bd.mat];
    PointProps pt;
    pt.e = pt.e_n = md.e;
    pt.nu = pt.nu_n = md.nu;
    pt.hprime = md.hprime;
    pt.yld = md.yld;
    if (M.temps || M.curves) {
      const ThermDesc* td = M.therm + bd.mat;
      double gp_t = 0.0, gp_dt = 0.0;
      if (M.temps) point_temps<NN>(M, bd, gt, wi.first + e, g, gp_t, gp_dt);
      point_props(md, td, gp_t, gp_dt, pt);
      if (M.temps && M.temps_incr)
        thermal_strain(pt.alpha, pt.alpha_n, td->alpha_zero, gp_t, gp_dt,
                       deps);
    }
    if (MODE == 1) {
      double* ud = M.uddt + 6 * bd.q_off + q;
#pragma unroll
      for (int k = 0; k < 6; ++k) ud[k * SG] = deps[k];
    }
    double cgn[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) cgn[k] = sn[k * SG];
    if (MODE == 1) {
      // the material update is the mm10 kernel's
    } else if (nonlinear) {
      double h[kMM01Hist];
      double* hn = M.hist_n + bd.hist_off + q;
      double* hn1 = M.hist_n1 + bd.hist_off + q;
      double f11 = 0.0;
      if (step != 1) {
#pragma unroll
        for (int k = 0; k < kMM01Hist; ++k) h[k] = hn[k * SG];
        f11 = M.hist_n[bd.hist_off + 10 * SG + g];  // history(1,11)
      }
      // rplstr scatters history only for iter > 0
      if (!mm01_point(step, md, pt, cgn, deps, h, f11, sn1, hn1, cp, SG,
                      iter > 0, sig))
        atomicOr(&M.flags[0], 1);
    } else {
      // drive_01_update_a / _b: linear-elastic [D]
      const double one = 1.0, two = 2.0, e_ = pt.e, nu = pt.nu;
      double c1 = (e_ / ((one + nu) * (one - two * nu)));
      double c2 = (one - nu) * c1;
      double c3 = ((one - two * nu) / two) * c1;
      double c4 = nu * c1;
      const double dm[6][6] = {{c2, c4, c4, 0, 0, 0}, {c4, c2, c4, 0, 0, 0},
                               {c4, c4, c2, 0, 0, 0}, {0, 0, 0, c3, 0, 0},
                               {0, 0, 0, 0, c3, 0},   {0, 0, 0, 0, 0, c3}};
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        double sk = cgn[k];
#pragma unroll
        for (int m = 0; m < 6; ++m) sk = sk + dm[m][k] * deps[m];
        sig[k] = sk;
        sn1[k * SG] = sk;
      }
      sig[6] = sig[7] = 0.0;
      sn1[6 * SG] = sn1[7 * SG] = sn1[8 * SG] = 0.0;
#pragma unroll
      for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) cp[tri(r, c) * SG] = dm[r][c];
    }
  }
  if (MODE == 1) return;  // uniform across the workgroup
  __syncthreads();
  if (GEONL) {
    // rknifv_std: Jacobians and B-bar terms of the n+1 configuration
    if (active) jacobian(1.0, 1.0, nullptr);
    __syncthreads();
    if (BBAR) {
      vol_terms();
      __syncthreads();
    }
    if (active && iter > 0) {  // rplstr: [R n+1] kept for iter > 0
      const int64_t SG = (int64_t)bd.span * G;
      const int64_t q = (int64_t)(wi.first + e) * G + g;
      double* rr = M.rot_n1 + 9 * bd.q_off + q;
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) rr[(a + 3 * c) * SG] = rn1[3 * a + c];
    }
  }
  // increments no longer read: their LDS takes the stresses (the Cauchy
  // stresses for finite strains, gpifv1)
  if (active) {
    if (GEONL) {
      double q2[36], cs[6];
      geonl_q(rn1, 2, q2);
      geonl_qmul(q2, sig, cs);
#pragma unroll
      for (int k = 0; k < 6; ++k) s_sig[e][g][k] = cs[k];
    } else {
#pragma unroll
      for (int k = 0; k < 6; ++k) s_sig[e][g][k] = sig[k];
    }
    if (want_energy) {  // rstgp1_b terms, point-major like its loops
      double factor = GEONL ? dfn1 * dj0 * 1.0 * s_w[g]
                            : s_gam[e][g][9] * 1.0 * s_w[g];
      s_en0[g * cnt + e] = sig[6] * factor;
      s_en1[g * cnt + e] = sig[7] * factor;
    }
  }
  __syncthreads();

  // internal energy and plastic work of this chunk: a fixed-order sum
  // (4 consecutive terms per lane, then a butterfly over the wave)
  if (want_energy && tid < 64) {
    const int nt = cnt * G;
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int t = 4 * tid + k;
      if (t < nt) {
        a = a + s_en0[t];
        b = b + s_en1[t];
      }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      a = a + __shfl_xor(a, off);
      b = b + __shfl_xor(b, off);
    }
    if (tid == 0) {
      en_wg[kBlkSums * (slot0 + blockIdx.x)] = a;
      en_wg[kBlkSums * (slot0 + blockIdx.x) + 1] = b;
    }
  }

  // phase 4: element internal forces f = sum_gp B^T sigma w |J| (gpifv1),
  // stored node-major inside the element: word 3 j + c is the force of
  // component c at local node j (the reference's dof c*NN + j)
  const double scalar = bd.etype == W3D_TET10 ? 1.0 / 6.0 : 1.0;
  double fabs_sum = 0.0;
  // lanes = (element, node): the node's three force components
  for (int k = tid; k < cnt * NN; k += 256) {
    const int ee = k / NN, j = k % NN;
    double f[3] = {0.0, 0.0, 0.0};
    for (int gg = 0; gg < G; ++gg) {
      const double* ga = s_gam[ee][gg];
      double nx = s_dn[0][gg][j], ny = s_dn[1][gg][j], nz = s_dn[2][gg][j];
      double bx = ga[0] * nx + ga[1] * ny + ga[2] * nz;
      double by = ga[3] * nx + ga[4] * ny + ga[5] * nz;
      double bz = ga[6] * nx + ga[7] * ny + ga[8] * nz;
      const double* s = s_sig[ee][gg];
      double w = s_w[gg] * scalar, d = ga[9];
      {  // component x
        double b0 = bx, b1 = 0.0;
        if (BBAR) {
          double vx = s_vol[ee][3 * j];
          b0 = alpha * bx + beta_b * vx;
          b1 = beta_b * (vx - bx);
        }
        f[0] = f[0] + b0 * s[0] * w * d + b1 * s[1] * w * d +
               b1 * s[2] * w * d + by * s[3] * w * d + bz * s[5] * w * d;
      }
      {  // component y
        double b1 = by, b0 = 0.0;
        if (BBAR) {
          double vy = s_vol[ee][3 * j + 1];
          b0 = beta_b * (vy - by);
          b1 = alpha * by + beta_b * vy;
        }
        f[1] = f[1] + b0 * s[0] * w * d + b1 * s[1] * w * d +
               b0 * s[2] * w * d + bx * s[3] * w * d + bz * s[4] * w * d;
      }
      {  // component z
        double b2 = bz, b0 = 0.0;
        if (BBAR) {
          double vz = s_vol[ee][3 * j + 2];
          b0 = beta_b * (vz - bz);
          b2 = alpha * bz + beta_b * vz;
        }
        f[2] = f[2] + b0 * s[0] * w * d + b0 * s[1] * w * d +
               b2 * s[2] * w * d + by * s[4] * w * d + bx * s[5] * w * d;
      }
    }
    // rknifv: forces of a skewed node to constraint-compatible
    // coordinates (trnvecs, switch 1)
    if (bd.trn) {
      const int node = M.conn[bd.conn_off + (int64_t)(wi.first + ee) * NN + j];
      const int ti = M.trn_idx[node];
      if (ti >= 0) trn_apply(M.trn_mat + 9 * ti, f);
    }
    double* out = M.einfvec + bd.ifv_off + (int64_t)(wi.first + ee) * T + 3 * j;
    out[0] = f[0];
    out[1] = f[1];
    out[2] = f[2];
    fabs_sum = fabs_sum + fabs(f[0]);
    fabs_sum = fabs_sum + fabs(f[1]);
    fabs_sum = fabs_sum + fabs(f[2]);
  }
  if (want_energy) {  // uniform: addifv's sum of |element force| terms
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1)
      fabs_sum = fabs_sum + __shfl_xor(fabs_sum, off);
    if ((tid & 63) == 0) s_red[tid >> 6] = fabs_sum;
    __syncthreads();
    if (tid == 0)
      en_wg[kBlkSums * (slot0 + blockIdx.x) + 2] =
          ((s_red[0] + s_red[1]) + s_red[2]) + s_red[3];
  }
  for (int k = tid; k < cnt; k += 256) {
    double v = 0.0;
    for (int gg = 0; gg < G; ++gg) v = v + (s_w[gg] * scalar) * s_gam[k][gg][9];
    M.evol[(int64_t)bd.felem - 1 + wi.first + k] = v;
  }
}

// ---------------------------------------------------------------------
// Gathers (deterministic: each output sums its inputs in a fixed order)
// ---------------------------------------------------------------------
__global__ void k_ifv_gather(const double* __restrict__ einfvec,
                             const int64_t* __restrict__ inv_ptr,
                             const int64_t* __restrict__ inv_off, int nonode,
                             double* __restrict__ ifv) {
  const int n = (int)xcd_block(blockIdx.x, gridDim.x) * blockDim.x +
                threadIdx.x;
  if (n >= nonode) return;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  for (int64_t p = inv_ptr[n]; p < inv_ptr[n + 1]; ++p) {
    int64_t o = inv_off[p];  // the node's 3 contiguous force words
    fx = fx + einfvec[o];
    fy = fy + einfvec[o + 1];
    fz = fz + einfvec[o + 2];
  }
  ifv[3 * n] = fx;
  ifv[3 * n + 1] = fy;
  ifv[3 * n + 2] = fz;
}

// Per-block internal energy, plastic work and sum of |element force|: the
// block's chunk sums in element order (one thread per block).
__global__ void k_energy_reduce(const double* __restrict__ en_wg,
                                const int32_t* __restrict__ wfirst,
                                const int32_t* __restrict__ wcount, int nblk,
                                double* __restrict__ sums) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  double en = 0.0, pw = 0.0, fa = 0.0;
  for (int w = wfirst[b]; w < wfirst[b] + wcount[b]; ++w) {
    en = en + en_wg[kBlkSums * w];
    pw = pw + en_wg[kBlkSums * w + 1];
    fa = fa + en_wg[kBlkSums * w + 2];
  }
  sums[kBlkSums * b] = en;
  sums[kBlkSums * b + 1] = pw;
  sums[kBlkSums * b + 2] = fa;
}

// Row-by-row assembly (assem_a_row, assemble_code.f:785) over the segment
// map: one thread per segment (a diagonal, or a run of one row's
// coefficients in the components of one column node).  The wave finds its
// segments' first contribution and first output by prefix sums of the
// headers from the wave's two base offsets.  Each output sums its
// contributions in FP64 in the reference's order (elements on the row's
// node, then erow, then ecol occurrences).  Outputs 0 .. neq-1 are the
// diagonal, the rest the coefficients right of the diagonal.
__device__ inline int wave_incl_scan(int x, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

template <typename KT>
__global__ void __launch_bounds__(256)
    k_assemble(const KT* __restrict__ ek, const uint16_t* __restrict__ hdr,
               const uint32_t* __restrict__ src,
               const int64_t* __restrict__ wbase, int64_t nseg, int64_t neq,
               double* __restrict__ diag, double* __restrict__ coef) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int h = t < nseg ? hdr[t] : 0;
  const int n = h & 255, mask = (h >> 8) & 7;
  const int len = __popc(mask);
  const int n_ex = wave_incl_scan(n, lane) - n;
  const int l_ex = wave_incl_scan(len, lane) - len;
  if (t >= nseg) return;
  const int64_t w = t >> 6;
  const int64_t so = wbase[2 * w] + n_ex;
  int64_t oo = wbase[2 * w + 1] + l_ex;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int k = 0; k < n; ++k) {
    const uint32_t b = src[so + k];
    if (mask & 1) s0 = s0 + (double)ek[b];
    if (mask & 2) s1 = s1 + (double)ek[b + 1];
    if (mask & 4) s2 = s2 + (double)ek[b + 2];
  }
  const double v[3] = {s0, s1, s2};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (!(mask >> c & 1)) continue;
    if (oo < neq)
      diag[oo] = v[c];
    else
      coef[oo - neq] = v[c];
    ++oo;
  }
}

// ---------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------
int group_key(int nnode, int ngp, int bbar, int mm10, int geonl, int trn) {
  return ((((nnode * 100 + ngp) * 2 + (bbar ? 1 : 0)) * 2 + (trn ? 1 : 0)) *
              2 + (mm10 ? 1 : 0)) *
             2 + (geonl ? 1 : 0);
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
                   int nonlinear, const double* d_u, const double* d_du,
                   int want_energy, double* en_wg, int pass, hipStream_t s) {
  (void)hblocks;
  for (size_t k = 0; k < groups.size(); ++k) {
    if (gcount[k] == 0) continue;
    const bool mm10 = (groups[k] >> 1) & 1;
    const bool geonl = groups[k] & 1;
    if (pass == 1 && !mm10) continue;
    const int mode = mm10 ? (pass == 0 ? 1 : 2) : 0;
    const WorkItem* w = d_work + gstart[k];
    bool done = false;
#define X(NN, G, BB)                                                        \
  if (!done && (groups[k] >> 3) == (group_key(NN, G, BB) >> 3)) {          \
    if (geonl && mode == 0)                                                 \
      hipLaunchKernelGGL((k_recover<NN, G, BB, 0, true>), dim3(gcount[k]),  \
                         dim3(256), 0, s, M, w, step, iter, nonlinear, d_u,  \
                         d_du, want_energy, en_wg, gstart[k]);              \
    else if (geonl && mode == 1)                                            \
      hipLaunchKernelGGL((k_recover<NN, G, BB, 1, true>), dim3(gcount[k]),  \
                         dim3(256), 0, s, M, w, step, iter, nonlinear, d_u,  \
                         d_du, want_energy, en_wg, gstart[k]);              \
    else if (geonl)                                                         \
      hipLaunchKernelGGL((k_recover<NN, G, BB, 2, true>), dim3(gcount[k]),  \
                         dim3(256), 0, s, M, w, step, iter, nonlinear, d_u,  \
                         d_du, want_energy, en_wg, gstart[k]);              \
    else if (mode == 0)                                                     \
      hipLaunchKernelGGL((k_recover<NN, G, BB, 0>), dim3(gcount[k]),        \
                         dim3(256), 0, s, M, w, step, iter, nonlinear, d_u,  \
                         d_du, want_energy, en_wg, gstart[k]);              \
    else if (mode == 1)                                                     \
      hipLaunchKernelGGL((k_recover<NN, G, BB, 1>), dim3(gcount[k]),        \
                         dim3(256), 0, s, M, w, step, iter, nonlinear, d_u,  \
                         d_du, want_energy, en_wg, gstart[k]);              \
    else                                                                    \
      hipLaunchKernelGGL((k_recover<NN, G, BB, 2>), dim3(gcount[k]),        \
                         dim3(256), 0, s, M, w, step, iter, nonlinear, d_u,  \
                         d_du, want_energy, en_wg, gstart[k]);              \
    done = true;                                                            \
  }
    W3D_GROUPS(X)
#undef X
    if (!done) return W3D_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

__global__ void k_fill_i32(int32_t* __restrict__ d, int n, int32_t v) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) d[k] = v;
}

int launch_fill_i32(int32_t* d, int n, int32_t v, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, d,
                       n, v);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_ifv_gather(const double* einfvec, const int64_t* inv_ptr,
                      const int64_t* inv_off, int nonode, double* ifv,
                      hipStream_t s) {
  int nb = (nonode + 255) / 256;
  if (nb > 0)
    hipLaunchKernelGGL(k_ifv_gather, dim3(nb), dim3(256), 0, s, einfvec,
                       inv_ptr, inv_off, nonode, ifv);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_energy_reduce(const double* en_wg, const int32_t* wfirst,
                         const int32_t* wcount, int nblk, double* sums,
                         hipStream_t s) {
  int nb = (nblk + 63) / 64;
  if (nb > 0)
    hipLaunchKernelGGL(k_energy_reduce, dim3(nb), dim3(64), 0, s, en_wg,
                       wfirst, wcount, nblk, sums);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_assemble(const double* ek64, const float* ek32, int fp32,
                    const uint16_t* hdr, const uint32_t* src,
                    const int64_t* wbase, int64_t nseg, int64_t neq,
                    double* diag, double* coef, hipStream_t s) {
  const int64_t nb = (nseg + 255) / 256;
  if (nb > 0) {
    if (fp32)
      hipLaunchKernelGGL(k_assemble<float>, dim3((unsigned)nb), dim3(256), 0,
                         s, ek32, hdr, src, wbase, nseg, neq, diag, coef);
    else
      hipLaunchKernelGGL(k_assemble<double>, dim3((unsigned)nb), dim3(256), 0,
                         s, ek64, hdr, src, wbase, nseg, neq, diag, coef);
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

}  // namespace w3d
