// w3d_asym.hip -- asymmetric assembly: the full element matrices and their
// gather into full CSR rows, FP64.
//
// Reference: gptns1.f:1513 bdbt_asym (BD(j,k) = sum_m D(m,k) B(m,j), then
// ek_full(j) += sum_r B(r,col) BD(row,r) for j = (col-1) T + row, every
// row and column), gptns1.f:1371 drive_10_cnst ([D] = the point's history
// words 1-36, column-major, times w |J|), gptns1.f:1174 extract_D_symmetric
// (mm01), gpmas1.f:399 adjust_cnst (1/6 for tets), assemble_code.f:1050
// assem_by_row_asymmetric / assem_a_row_asym (row srow gathers the words
// (erow-1) T + ecol of each element on its node, elements in
// inverse-incidence order, the column found by binary search).
//
// One 256-lane workgroup per element: the point loop stages B and [D] B of
// one point in LDS and every lane accumulates a 4x4 (or 2x2) block of the
// T*T words
// in registers.  The assembly runs one wave per equation row over a map
// built once per pattern: for each (node, element) incidence, the Ke base of
// the element and, per element column, its position in the CSR rows of the
// node (the same for the node's three rows), or none if constrained.
#include <hip/hip_runtime.h>

#include <cmath>

#include "w3d_internal.h"

namespace w3d {

template <int NN, int G, bool BBAR>
__global__ void __launch_bounds__(256)
    k_tangent_full(DevModel M, const WorkItem* __restrict__ work,
                   const int64_t* __restrict__ ekf_off,
                   double* __restrict__ ekf) {
  constexpr int T = 3 * NN;
  // each lane owns an RB x RB block of (column, row) words: RB values of
  // B and RB of [D] B per strain component feed RB^2 products
  constexpr int RB = (T % 4 == 0 && (T / 4) * (T / 4) <= 256 && T >= 48) ? 4 : 2;
  constexpr int NB = T / RB;
  static_assert(T % RB == 0 && NB * NB <= 256, "register blocking");
  __shared__ double s_ce[T];
  __shared__ double s_g[G][9];
  __shared__ double s_dj[G];
  __shared__ double s_vol[24];
  __shared__ double s_b[6 * T];
  __shared__ double s_bd[6 * T];
  __shared__ double s_d[36];
  const WorkItem wi = work[blockIdx.x];
  const BlockDesc bd = M.blocks[wi.blk];
  const GaussTable* gt = M.tables + bd.table;
  const int tid = threadIdx.x;
  const int64_t SG = (int64_t)bd.span * G;
  const double scalar = bd.etype == W3D_TET10 ? 1.0 / 6.0 : 1.0;
  const double one = 1.0, two = 2.0, third = 1.0 / 3.0;
  const double alpha = (two + M.eps_bbar) * third;
  const double beta = (one - M.eps_bbar) * third;
  for (int e = 0; e < wi.count; ++e) {
    const int i = wi.first + e;
    __syncthreads();
    if (tid < NN) {
      const int node = M.conn[bd.conn_off + (int64_t)i * NN + tid];
      s_ce[tid] = M.coords[3 * node];
      s_ce[NN + tid] = M.coords[3 * node + 1];
      s_ce[2 * NN + tid] = M.coords[3 * node + 2];
    }
    __syncthreads();
    if (tid < G) {  // jacob1
      const double* nxi = gt->dn[0][tid];
      const double* neta = gt->dn[1][tid];
      const double* nzeta = gt->dn[2][tid];
      double jac[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) jac[k] = 0.0;
      for (int j = 0; j < NN; ++j) {
        const double x = s_ce[j], y = s_ce[NN + j], z = s_ce[2 * NN + j];
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
      const double j1 = jac[4] * jac[8] - jac[5] * jac[7];
      const double j2 = jac[3] * jac[8] - jac[5] * jac[6];
      const double j3 = jac[3] * jac[7] - jac[4] * jac[6];
      const double d = jac[0] * j1 - jac[1] * j2 + jac[2] * j3;
      s_dj[tid] = d;
      double* g = s_g[tid];
      g[0] = j1 / d;
      g[3] = -j2 / d;
      g[6] = j3 / d;
      g[1] = (jac[7] * jac[2] - jac[1] * jac[8]) / d;
      g[4] = (jac[0] * jac[8] - jac[6] * jac[2]) / d;
      g[7] = (jac[1] * jac[6] - jac[0] * jac[7]) / d;
      g[2] = (jac[1] * jac[5] - jac[2] * jac[4]) / d;
      g[5] = (jac[2] * jac[3] - jac[0] * jac[5]) / d;
      g[8] = (jac[0] * jac[4] - jac[1] * jac[3]) / d;
    }
    __syncthreads();
    if (BBAR && tid < 24) {  // vol_terms, vol_avg
      const int k = tid % 3, j = tid / 3;
      double v = 0.0, volume = 0.0;
      for (int g = 0; g < G; ++g) {
        const double* a = s_g[g];
        v = v + (a[3 * k] * gt->dn[0][g][j] + a[3 * k + 1] * gt->dn[1][g][j] +
                 a[3 * k + 2] * gt->dn[2][g][j]) *
                    s_dj[g];
        volume = volume + s_dj[g];
      }
      s_vol[3 * j + k] = v * (1.0 / volume);
    }
    const bool own = tid < NB * NB;
    const int col0 = RB * (tid / NB), row0 = RB * (tid % NB);
    double acc[RB][RB];
#pragma unroll
    for (int a = 0; a < RB; ++a)
#pragma unroll
      for (int c = 0; c < RB; ++c) acc[a][c] = 0.0;
    for (int g = 0; g < G; ++g) {
      __syncthreads();
      for (int k = tid; k < 6 * T; k += 256) s_b[k] = 0.0;
      __syncthreads();
      if (tid < NN) {  // the B columns of node j (blcmp1)
        const int j = tid;
        const double* a = s_g[g];
        const double nx = gt->dn[0][g][j], ny = gt->dn[1][g][j],
                     nz = gt->dn[2][g][j];
        const double bx = a[0] * nx + a[1] * ny + a[2] * nz;
        const double by = a[3] * nx + a[4] * ny + a[5] * nz;
        const double bz = a[6] * nx + a[7] * ny + a[8] * nz;
        s_b[0 * T + j] = bx;
        s_b[3 * T + j] = by;
        s_b[5 * T + j] = bz;
        s_b[1 * T + NN + j] = by;
        s_b[3 * T + NN + j] = bx;
        s_b[4 * T + NN + j] = bz;
        s_b[2 * T + 2 * NN + j] = bz;
        s_b[4 * T + 2 * NN + j] = by;
        s_b[5 * T + 2 * NN + j] = bx;
      }
      if (tid >= 64 && tid < 100) {  // [D] of the point times w |J|
        const int t = tid - 64, m = t % 6, kk = t / 6;
        const double f = gt->w[g] * s_dj[g];
        const int64_t q = (int64_t)i * G + g;
        double v;
        if (bd.model == W3D_MM10) {
          v = M.hist_n1[bd.hist_off + (int64_t)(m + 6 * kk) * SG + q] * f;
        } else {
          const int r = m > kk ? m : kk, c = m > kk ? kk : m;
          v = f * M.cep[21 * bd.q_off + (int64_t)(r * (r + 1) / 2 + c) * SG +
                        q];
        }
        if (scalar != 1.0) v = scalar * v;
        s_d[6 * m + kk] = v;
      }
      __syncthreads();
      if (BBAR && tid < 8) {  // bmod, node j
        const int j = tid;
        double* r1 = s_b;
        double* r2 = s_b + T;
        double* r3 = s_b + 2 * T;
        const double d = r1[j];
        r1[j] = alpha * d + beta * s_vol[3 * j + 0];
        r2[j] = beta * (s_vol[3 * j + 0] - d);
        r1[8 + j] = beta * (s_vol[3 * j + 1] - r2[8 + j]);
        r1[16 + j] = beta * (s_vol[3 * j + 2] - r3[16 + j]);
        r2[8 + j] = alpha * r2[8 + j] + beta * s_vol[3 * j + 1];
        r2[16 + j] = r1[16 + j];
        r3[j] = r2[j];
        r3[8 + j] = r1[8 + j];
        r3[16 + j] = alpha * r3[16 + j] + beta * s_vol[3 * j + 2];
      }
      __syncthreads();
      for (int k = tid; k < 6 * T; k += 256) {
        const int kk = k / T, j = k % T;
        s_bd[k] = s_d[0 * 6 + kk] * s_b[0 * T + j] +
                  s_d[1 * 6 + kk] * s_b[1 * T + j] +
                  s_d[2 * 6 + kk] * s_b[2 * T + j] +
                  s_d[3 * 6 + kk] * s_b[3 * T + j] +
                  s_d[4 * 6 + kk] * s_b[4 * T + j] +
                  s_d[5 * 6 + kk] * s_b[5 * T + j];
      }
      __syncthreads();
      if (own) {
        // word (col, row) += sum_r B(r, col) BD(r, row), r in order
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          double bc[RB], br[RB];
#pragma unroll
          for (int a = 0; a < RB; ++a) {
            bc[a] = s_b[r * T + col0 + a];
            br[a] = s_bd[r * T + row0 + a];
          }
#pragma unroll
          for (int a = 0; a < RB; ++a)
#pragma unroll
            for (int c = 0; c < RB; ++c) acc[a][c] = acc[a][c] + bc[a] * br[c];
        }
      }
    }
    double* out = ekf + ekf_off[wi.blk] + (int64_t)i * T * T;
    if (own) {
#pragma unroll
      for (int a = 0; a < RB; ++a)
#pragma unroll
        for (int c = 0; c < RB; ++c) out[(col0 + a) * T + row0 + c] = acc[a][c];
    }
  }
}

// One wave per equation row: the row's values are summed in LDS, element
// after element in inverse-incidence order, lane ecol adding the element's
// word (erow-1) T + ecol (distinct columns within an element, so no two
// lanes meet); the Ke words of an element row and the row's values are
// read and written contiguously.
constexpr int kAsmWaves = 4;
constexpr int kAsmChunk = 8;  // elements whose loads are in flight together
__device__ inline void wave_fence() { __builtin_amdgcn_wave_barrier(); }

__global__ void __launch_bounds__(64 * kAsmWaves)
    k_assemble_full(const double* __restrict__ ekf,
                    const int64_t* __restrict__ ebase,
                    const int64_t* __restrict__ atoff,
                    const uint16_t* __restrict__ at,
                    const int32_t* __restrict__ eq_dof,
                    const int64_t* __restrict__ inv_ptr,
                    const int32_t* __restrict__ inv_ln,
                    const int64_t* __restrict__ rp, double* __restrict__ val,
                    int neq) {
  __shared__ double s_row[kAsmWaves][kAsmMaxRow];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = blockIdx.x * kAsmWaves + wv;
  if (r >= neq) return;  // uniform across the wave
  const int64_t r0 = rp[r], r1 = rp[r + 1];
  const int len = (int)(r1 - r0);  // <= kAsmMaxRow (checked on the host)
  double* row = s_row[wv];
  for (int k = lane; k < len; k += 64) row[k] = 0.0;
  wave_fence();
  const int dof = eq_dof[r], node = dof / 3, comp = dof % 3;
  const int64_t pb = inv_ptr[node], pe = inv_ptr[node + 1];
  // elements in chunks of kAsmChunk: their loads are independent and
  // overlap; the sums then go into LDS in element order
  for (int64_t p0 = pb; p0 < pe; p0 += kAsmChunk) {
    int pos[kAsmChunk];
    double v[kAsmChunk];
#pragma unroll
    for (int j = 0; j < kAsmChunk; ++j) {
      pos[j] = kAsmNone;
      v[j] = 0.0;
      const int64_t p = p0 + j;
      if (p >= pe) continue;
      const int64_t a0 = atoff[p];
      const int T = (int)(atoff[p + 1] - a0);
      if (lane >= T) continue;
      pos[j] = at[a0 + lane];
      if (pos[j] == kAsmNone) continue;
      const int erow = comp * (T / 3) + inv_ln[p];
      v[j] = ekf[ebase[p] + erow * T + lane];
    }
#pragma unroll
    for (int j = 0; j < kAsmChunk; ++j) {
      if (pos[j] != kAsmNone) row[pos[j]] = row[pos[j]] + v[j];
      wave_fence();
    }
  }
  for (int k = lane; k < len; k += 64) val[r0 + k] = row[k];
}

// ---------------------------------------------------------------------
// Jacobi-preconditioned BiCGStab (van der Vorst 1992, right
// preconditioning) on the full CSR rows, FP64.  The scalars live in
// device memory (sc: rho, rho_prev, alpha, omega, (rh, v), (t, s), (t, t),
// (r, r)) so an iteration is a fixed sequence of launches; dot products
// use the solver's fixed reduction tree, so a solve is bitwise repeatable.
// ---------------------------------------------------------------------
__global__ void k_csr_diag_inv(const int64_t* __restrict__ rp,
                               const int32_t* __restrict__ col,
                               const double* __restrict__ val, int n,
                               double* __restrict__ dinv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int64_t k = rp[r]; k < rp[r + 1]; ++k)
    if (col[k] == r) d = val[k];
  dinv[r] = d != 0.0 ? 1.0 / d : 1.0;
}

// The scalars of one BiCGStab iteration live in sc (device): [0] rho =
// (rh, r), [1] rho of the last iteration, [2] alpha of the last, [3] omega
// of the last, [4] (rh, v), [5] (t, s), [6] (t, t), [7] (r, r), [8] the
// breakdown flag, [9] beta, [10] alpha, [11] omega.  One thread forms them
// between the dot products; a breakdown (a zero rho, (rh, v), last rho or
// last omega) raises the flag, after which no vector kernel changes x or r,
// so x keeps the last iterate.
__global__ void k_bicg_beta(double* __restrict__ sc, int first) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (sc[0] == 0.0 || (!first && (sc[1] == 0.0 || sc[3] == 0.0))) {
    sc[8] = 1.0;
    sc[9] = 0.0;
    return;
  }
  sc[9] = first ? 0.0 : (sc[0] / sc[1]) * (sc[2] / sc[3]);
}

__global__ void k_bicg_alpha(double* __restrict__ sc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (sc[8] != 0.0 || sc[4] == 0.0) {
    sc[8] = 1.0;
    sc[10] = 0.0;
    return;
  }
  sc[10] = sc[0] / sc[4];
}

// omega = (t, s) / (t, t); (t, t) = 0: s is already the residual of x (a
// converged half step), omega = 0
__global__ void k_bicg_omega(double* __restrict__ sc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sc[11] = sc[6] != 0.0 ? sc[5] / sc[6] : 0.0;
}

// p = r + beta (p - omega_last v), y = p / diag
__global__ void k_bicg_p(int n, const double* __restrict__ sc, int first,
                         const double* __restrict__ r,
                         const double* __restrict__ v,
                         const double* __restrict__ dinv,
                         double* __restrict__ p, double* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || sc[8] != 0.0) return;
  double pi;
  if (first)
    pi = r[i];
  else
    pi = r[i] + sc[9] * (p[i] - sc[3] * v[i]);
  p[i] = pi;
  y[i] = pi * dinv[i];
}

// s = r - alpha v; z = s / diag (x takes alpha y with omega z in k_bicg_x)
__global__ void k_bicg_s(int n, const double* __restrict__ sc,
                         const double* __restrict__ r,
                         const double* __restrict__ v,
                         const double* __restrict__ dinv,
                         double* __restrict__ s, double* __restrict__ z) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || sc[8] != 0.0) return;
  const double si = r[i] - sc[10] * v[i];
  s[i] = si;
  z[i] = si * dinv[i];
}

// x += alpha y + omega z (alpha y first); r = s - omega t
__global__ void k_bicg_x(int n, const double* __restrict__ sc,
                         const double* __restrict__ s,
                         const double* __restrict__ t,
                         const double* __restrict__ y,
                         const double* __restrict__ z,
                         double* __restrict__ x, double* __restrict__ r) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || sc[8] != 0.0) return;
  const double omega = sc[11];
  x[i] = (x[i] + sc[10] * y[i]) + omega * z[i];
  r[i] = s[i] - omega * t[i];
}

// the next iteration's rho_last, alpha_last, omega_last
__global__ void k_bicg_shift(double* __restrict__ sc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  sc[1] = sc[0];
  sc[2] = sc[10];
  sc[3] = sc[11];
}

__global__ void k_fill_f64(double* __restrict__ d, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = v;
}

#define ACHK(x)                                  \
  do {                                           \
    if ((x) != hipSuccess) return W3D_ERR_DEVICE; \
  } while (0)

int bicgstab_words(int n) { return 9 * n + kBicgScalars + solver_npart(n); }

int bicgstab(int n, const int64_t* rp, const int32_t* col, const double* val,
             const double* b, double* x, double* work, double rtol, int maxit,
             int* iters, double* relres, hipStream_t st) {
  double* r = work;
  double* rh = r + n;
  double* p = rh + n;
  double* v = p + n;
  double* s = v + n;
  double* t = s + n;
  double* y = t + n;
  double* z = y + n;
  double* dinv = z + n;
  double* sc = dinv + n;
  double* part = sc + kBicgScalars;
  const dim3 gb((n + 255) / 256), tb(256);
  // fills by kernels, not memsets: the work array may be a static device
  // array (w3d_solve_bicgstab_noalloc), and no memset node enters a graph
  hipLaunchKernelGGL(k_csr_diag_inv, gb, tb, 0, st, rp, col, val, n, dinv);
  ACHK(hipMemcpyAsync(r, b, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
  ACHK(hipMemcpyAsync(rh, b, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_fill_f64, gb, tb, 0, st, x, n, 0.0);
  hipLaunchKernelGGL(k_fill_f64, gb, tb, 0, st, v, n, 0.0);
  hipLaunchKernelGGL(k_fill_f64, dim3(1), dim3(64), 0, st, sc, kBicgScalars,
                     0.0);
  int rc;
  if ((rc = solver_dot(n, b, b, part, sc + 7, st))) return rc;
  double bb = 0.0;
  ACHK(hipMemcpyAsync(&bb, sc + 7, 8, hipMemcpyDeviceToHost, st));
  ACHK(hipStreamSynchronize(st));
  const double bnorm = std::sqrt(bb);
  *iters = 0;
  *relres = 0.0;
  if (bnorm == 0.0) return W3D_OK;
  for (int it = 1; it <= maxit; ++it) {
    const int first = it == 1 ? 1 : 0;
    if ((rc = solver_dot(n, rh, r, part, sc + 0, st))) return rc;
    hipLaunchKernelGGL(k_bicg_beta, dim3(1), dim3(64), 0, st, sc, first);
    hipLaunchKernelGGL(k_bicg_p, gb, tb, 0, st, n, sc, first, r, v, dinv, p,
                       y);
    if ((rc = solver_spmv(n, rp, col, val, y, v, st))) return rc;
    if ((rc = solver_dot(n, rh, v, part, sc + 4, st))) return rc;
    hipLaunchKernelGGL(k_bicg_alpha, dim3(1), dim3(64), 0, st, sc);
    hipLaunchKernelGGL(k_bicg_s, gb, tb, 0, st, n, sc, r, v, dinv, s, z);
    if ((rc = solver_spmv(n, rp, col, val, z, t, st))) return rc;
    if ((rc = solver_dot(n, t, s, part, sc + 5, st))) return rc;
    if ((rc = solver_dot(n, t, t, part, sc + 6, st))) return rc;
    hipLaunchKernelGGL(k_bicg_omega, dim3(1), dim3(64), 0, st, sc);
    hipLaunchKernelGGL(k_bicg_x, gb, tb, 0, st, n, sc, s, t, y, z, x, r);
    hipLaunchKernelGGL(k_bicg_shift, dim3(1), dim3(64), 0, st, sc);
    if ((rc = solver_dot(n, r, r, part, sc + 7, st))) return rc;
    double rr[2] = {0.0, 0.0};
    ACHK(hipMemcpyAsync(rr, sc + 7, 16, hipMemcpyDeviceToHost, st));
    ACHK(hipStreamSynchronize(st));
    *relres = std::sqrt(rr[0]) / bnorm;
    if (rr[1] != 0.0) return W3D_ERR_NUMERIC;  // breakdown: x as it was
    *iters = it;
    if (!std::isfinite(rr[0])) return W3D_ERR_NUMERIC;
    if (*relres <= rtol) return W3D_OK;
  }
  return W3D_ERR_NUMERIC;
}

#undef ACHK

#define W3D_FULL_GROUPS(X) \
  X(8, 8, true)            \
  X(8, 8, false)           \
  X(20, 27, false)         \
  X(20, 8, false)          \
  X(20, 14, false)         \
  X(20, 9, false)          \
  X(10, 4, false)          \
  X(10, 1, false)          \
  X(10, 5, false)

int launch_tangent_full(const DevModel& M, const WorkItem* d_work,
                        const std::vector<int>& groups,
                        const std::vector<int>& gstart,
                        const std::vector<int>& gcount,
                        const int64_t* d_ekf_off, double* d_ekf,
                        hipStream_t s) {
  for (size_t k = 0; k < groups.size(); ++k) {
    if (gcount[k] == 0) continue;
    if (groups[k] & 1) return W3D_ERR_UNSUPPORTED;  // finite strains
    const WorkItem* w = d_work + gstart[k];
    bool done = false;
#define X(NN, G, BB)                                                      \
  if (!done && (groups[k] >> 3) == (group_key(NN, G, BB, 0, 0) >> 3)) {  \
    hipLaunchKernelGGL((k_tangent_full<NN, G, BB>), dim3(gcount[k]),      \
                       dim3(256), 0, s, M, w, d_ekf_off, d_ekf);          \
    done = true;                                                          \
  }
    W3D_FULL_GROUPS(X)
#undef X
    if (!done) return W3D_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_assemble_full(const double* ekf, const int64_t* ebase,
                         const int64_t* atoff, const uint16_t* at,
                         const int32_t* eq_dof, const int64_t* inv_ptr,
                         const int32_t* inv_ln, const int64_t* rp,
                         double* val, int neq, hipStream_t s) {
  if (neq == 0) return W3D_OK;
  hipLaunchKernelGGL(k_assemble_full,
                     dim3((neq + kAsmWaves - 1) / kAsmWaves),
                     dim3(64 * kAsmWaves), 0, s, ekf, ebase, atoff, at,
                     eq_dof, inv_ptr, inv_ln, rp, val, neq);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

}  // namespace w3d
