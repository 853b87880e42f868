// w3d_solver.hip -- device Jacobi-preconditioned conjugate gradients on the
// assembled stiffness (the upper-triangular VSS arrays expanded to a full
// symmetric CSR), FP64 throughout.
//
// Stands in for the reference's direct solve of the assembled equations
// (drive_assemble_solve.f:654 -> drive_pardiso.f:11, and the MKL iterative
// option of the same driver); the reference's element-by-element PCG
// (old_code/EBE_solver/lnpcg_ser/lnpcg.f) is the precedent for a Jacobi
// preconditioner.  Every reduction is a fixed tree (per-block partial sums
// then one block over the partials), so a solve is bitwise repeatable.
#include <hip/hip_runtime.h>

#include <mutex>

#include "w3d_internal.h"

namespace w3d {

constexpr int kRedBlock = 256;

// Work memory of w3d_solve_pcg_static, which allocates nothing: the vectors
// r, z, p, q of up to kStaticMaxEq equations, then 8 scalars and the dot
// products' partial sums (solver_npart <= 1024).  Each device holds its own
// copy for the life of the process; concurrent calls on one device share
// it.
constexpr int64_t kStaticWords = 4LL * kStaticMaxEq + 8 + 1024;
static __device__ __align__(256) double scratch[kStaticWords];

// The device address of this device's scratch (hipGetSymbolAddress once
// per device, with that device current; the host-side name is not a
// device address).
int solver_static_scratch(int device, double** out) {
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return W3D_ERR_ARG;
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
      return W3D_ERR_DEVICE;
    addr[device] = static_cast<double*>(p);
  }
  *out = addr[device];
  return W3D_OK;
}

// vals[k] = the assembled word vidx[k] (diagonal words first, then the
// upper-triangle coefficients): the full CSR values of an assembly
__global__ void k_csr_values(const double* __restrict__ diag,
                             const double* __restrict__ coef, int neq,
                             const uint32_t* __restrict__ vidx, int64_t nnz,
                             double* __restrict__ vals) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) {
    uint32_t v = vidx[k];
    vals[k] = v < (uint32_t)neq ? diag[v] : coef[v - neq];
  }
}

// y = A x, 16 lanes per row; each lane sums its strided share of the row,
// then a fixed butterfly over the 16 lanes
__global__ void __launch_bounds__(256)
    k_spmv(int n, const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
           const double* __restrict__ val, const double* __restrict__ x,
           double* __restrict__ y) {
  const int lane = threadIdx.x & 15;
  const int row = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4);
  double s = 0.0;
  if (row < n) {
    for (int64_t k = rp[row] + lane; k < rp[row + 1]; k += 16)
      s = s + val[k] * x[col[k]];
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) s = s + __shfl_xor(s, off, 16);
  if (row < n && lane == 0) y[row] = s;
}

// block partial sums of a . b
__global__ void __launch_bounds__(kRedBlock)
    k_dot_partial(int n, const double* __restrict__ a,
                  const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double s[kRedBlock];
  double t = 0.0;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n;
       i += gridDim.x * kRedBlock)
    t = t + a[i] * b[i];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int w = kRedBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(kRedBlock)
    k_sum_partials(int np, const double* __restrict__ part,
                   double* __restrict__ out) {
  __shared__ double s[kRedBlock];
  double t = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedBlock) t = t + part[i];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int w = kRedBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

// z = r / diag
__global__ void k_jacobi(int n, const double* __restrict__ diag,
                         const double* __restrict__ r, double* __restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = r[i] / diag[i];
}

// p = z + (rz_new / rz_old) p   (beta from the device scalars)
__global__ void k_update_p(int n, const double* __restrict__ z,
                           const double* __restrict__ sc,
                           double* __restrict__ p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = z[i] + (sc[1] / sc[0]) * p[i];
}

// x += alpha p, r -= alpha q with alpha = rz / pq
__global__ void k_update_xr(int n, const double* __restrict__ p,
                            const double* __restrict__ q,
                            const double* __restrict__ sc,
                            double* __restrict__ x, double* __restrict__ r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double alpha = sc[1] / sc[2];
    x[i] = x[i] + alpha * p[i];
    r[i] = r[i] - alpha * q[i];
  }
}

// r = b - r (after r = A x)
__global__ void k_residual(int n, const double* __restrict__ b,
                           double* __restrict__ r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - r[i];
}

// sc[0] = sc[1] (keep the previous r.z)
__global__ void k_shift(double* sc) { sc[0] = sc[1]; }

// ---------------------------------------------------------------------
// Batched Jacobi-PCG over independent systems.  Tile t of a launch is
// tile t - tile0 of the system that owns it; each system's reductions run
// over its own tiles in a fixed order, so a system's iterates do not
// depend on the other systems of the batch.
// ---------------------------------------------------------------------
constexpr int kPartBB = 0, kPartPQ = 1, kPartRR = 3, kPartRZ = 5;
constexpr int kSpmvLanes = 16;

// the system owning global tile t (tile0 ascending, system 0 at tile 0)
template <typename S>
__device__ __forceinline__ int owner(const S* __restrict__ sys, int nsys,
                                     int t) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (sys[mid].tile0 <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// sum of v over the workgroup (256 threads), the same order every time;
// every thread gets the result
__device__ __forceinline__ double block_sum(double v, double* s_red) {
  s_red[threadIdx.x] = v;
  __syncthreads();
  for (int w = kRedBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      s_red[threadIdx.x] = s_red[threadIdx.x] + s_red[threadIdx.x + w];
    __syncthreads();
  }
  const double r = s_red[0];
  __syncthreads();
  return r;
}

// sum of n partials (strided per thread, then the tree)
__device__ __forceinline__ double sum_partials(const double* __restrict__ v,
                                               int n, double* s_red) {
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += kRedBlock) t = t + v[i];
  return block_sum(t, s_red);
}

__global__ void __launch_bounds__(kRedBlock)
    k_pcg_init(const PcgSys* __restrict__ sys, int nsys,
               double* __restrict__ part) {
  __shared__ double s_red[kRedBlock];
  const int si = owner(sys, nsys, blockIdx.x);
  const PcgSys& S = sys[si];
  const int t = blockIdx.x - S.tile0;
  const int row = t * kTileRows + threadIdx.x;
  double bb = 0.0, rz = 0.0;
  if (threadIdx.x < kTileRows && row < S.n) {
    const double bi = S.b[row];
    const double zi = bi / S.diag[row];
    S.x[row] = 0.0;
    S.r[row] = bi;
    S.p[row] = zi;
    bb = bi * bi;
    rz = bi * zi;
  }
  bb = block_sum(bb, s_red);
  rz = block_sum(rz, s_red);
  if (threadIdx.x == 0) {
    double* P = part + 7 * (int64_t)S.tile0;
    P[kPartBB * S.ntile + t] = bb;
    P[(kPartRZ + 1) * S.ntile + t] = rz;  // the r.z of iteration -1
  }
}

// one workgroup per system: |b|, and a zero right-hand side is solved
__global__ void __launch_bounds__(kRedBlock)
    k_pcg_init2(const PcgSys* __restrict__ sys, const double* __restrict__ part,
                int32_t* __restrict__ state, double* __restrict__ sc) {
  __shared__ double s_red[kRedBlock];
  const PcgSys& S = sys[blockIdx.x];
  const double* P = part + 7 * (int64_t)S.tile0;
  const double bb = sum_partials(P + kPartBB * S.ntile, S.ntile, s_red);
  if (threadIdx.x == 0) {
    sc[2 * blockIdx.x] = sqrt(bb);
    sc[2 * blockIdx.x + 1] = bb;
    state[2 * blockIdx.x] = bb == 0.0 ? 1 : 0;
    state[2 * blockIdx.x + 1] = 0;
  }
}

// q = A p on the tile's rows (16 lanes per row) and the partial p.q
__global__ void __launch_bounds__(kRedBlock)
    k_pcg_spmv(const PcgSys* __restrict__ sys, int nsys, int it,
               double* __restrict__ part, const int32_t* __restrict__ state) {
  __shared__ double s_red[kRedBlock];
  const int si = owner(sys, nsys, blockIdx.x);
  if (state[2 * si]) return;
  const PcgSys& S = sys[si];
  const int t = blockIdx.x - S.tile0;
  const int lane = threadIdx.x % kSpmvLanes;
  constexpr int kRowsPerPass = kRedBlock / kSpmvLanes;
  double pq = 0.0;
  for (int k = threadIdx.x / kSpmvLanes; k < kTileRows; k += kRowsPerPass) {
    const int row = t * kTileRows + k;
    double v = 0.0;
    if (row < S.n) {
      for (int64_t j = S.rp[row] + lane; j < S.rp[row + 1]; j += kSpmvLanes)
        v = v + S.val[j] * S.p[S.col[j]];
    }
#pragma unroll
    for (int off = kSpmvLanes / 2; off >= 1; off >>= 1)
      v = v + __shfl_xor(v, off, kSpmvLanes);
    if (row < S.n && lane == 0) {
      S.q[row] = v;
      pq = pq + S.p[row] * v;
    }
  }
  pq = block_sum(pq, s_red);
  if (threadIdx.x == 0)
    part[7 * (int64_t)S.tile0 + (kPartPQ + (it & 1)) * S.ntile + t] = pq;
}

// alpha = r.z / p.q; x += alpha p, r -= alpha q; partial r.r and r.z
__global__ void __launch_bounds__(kRedBlock)
    k_pcg_xr(const PcgSys* __restrict__ sys, int nsys, int it,
             double* __restrict__ part, const int32_t* __restrict__ state) {
  __shared__ double s_red[kRedBlock];
  const int si = owner(sys, nsys, blockIdx.x);
  if (state[2 * si]) return;
  const PcgSys& S = sys[si];
  const int t = blockIdx.x - S.tile0;
  double* P = part + 7 * (int64_t)S.tile0;
  const double pq =
      sum_partials(P + (kPartPQ + (it & 1)) * S.ntile, S.ntile, s_red);
  const double rz =
      sum_partials(P + (kPartRZ + ((it - 1) & 1)) * S.ntile, S.ntile, s_red);
  const double alpha = rz / pq;
  const int row = t * kTileRows + threadIdx.x;
  double rr = 0.0, rzn = 0.0;
  if (threadIdx.x < kTileRows && row < S.n) {
    S.x[row] = S.x[row] + alpha * S.p[row];
    const double ri = S.r[row] - alpha * S.q[row];
    S.r[row] = ri;
    rr = ri * ri;
    rzn = ri * (ri / S.diag[row]);
  }
  rr = block_sum(rr, s_red);
  rzn = block_sum(rzn, s_red);
  if (threadIdx.x == 0) {
    P[(kPartRR + (it & 1)) * S.ntile + t] = rr;
    P[(kPartRZ + (it & 1)) * S.ntile + t] = rzn;
  }
}

// convergence |r| <= rtol |b| (or the iteration limit) ends the system;
// otherwise p = D^-1 r + (r.z / r.z old) p.  Every workgroup of a system
// reaches the same decision from the same partials; the first tile's
// writes the state, which a sibling in this launch may or may not see:
// either way it leaves p alone or updates a p that is no longer read.
__global__ void __launch_bounds__(kRedBlock)
    k_pcg_p(const PcgSys* __restrict__ sys, int nsys, int it,
            const double* __restrict__ part,
            int32_t* __restrict__ state, double* __restrict__ sc) {
  __shared__ double s_red[kRedBlock];
  const int si = owner(sys, nsys, blockIdx.x);
  if (state[2 * si]) return;
  const PcgSys& S = sys[si];
  const int t = blockIdx.x - S.tile0;
  const double* P = part + 7 * (int64_t)S.tile0;
  const double rr =
      sum_partials(P + (kPartRR + (it & 1)) * S.ntile, S.ntile, s_red);
  const bool conv = sqrt(rr) <= S.rtol * sc[2 * si];
  if (conv || it + 1 >= S.maxit) {
    if (t == 0 && threadIdx.x == 0) {
      state[2 * si] = conv ? 1 : 2;
      state[2 * si + 1] = it + 1;
      sc[2 * si + 1] = rr;
    }
    return;
  }
  const double rzn =
      sum_partials(P + (kPartRZ + (it & 1)) * S.ntile, S.ntile, s_red);
  const double rzo =
      sum_partials(P + (kPartRZ + ((it - 1) & 1)) * S.ntile, S.ntile, s_red);
  const double beta = rzn / rzo;
  const int row = t * kTileRows + threadIdx.x;
  if (threadIdx.x < kTileRows && row < S.n)
    S.p[row] = S.r[row] / S.diag[row] + beta * S.p[row];
}

int launch_pcg_batch_init(const PcgSys* d_sys, int nsys, int ntile_total,
                          double* part, int32_t* state, double* sc,
                          hipStream_t s) {
  if (nsys <= 0) return W3D_OK;
  if (ntile_total > 0)
    hipLaunchKernelGGL(k_pcg_init, dim3(ntile_total), dim3(kRedBlock), 0, s,
                       d_sys, nsys, part);
  hipLaunchKernelGGL(k_pcg_init2, dim3(nsys), dim3(kRedBlock), 0, s, d_sys,
                     part, state, sc);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_pcg_batch_iter(const PcgSys* d_sys, int nsys, int ntile_total,
                          int it, double* part, int32_t* state, double* sc,
                          hipStream_t s) {
  if (nsys <= 0 || ntile_total <= 0) return W3D_OK;
  hipLaunchKernelGGL(k_pcg_spmv, dim3(ntile_total), dim3(kRedBlock), 0, s,
                     d_sys, nsys, it, part, state);
  hipLaunchKernelGGL(k_pcg_xr, dim3(ntile_total), dim3(kRedBlock), 0, s,
                     d_sys, nsys, it, part, state);
  hipLaunchKernelGGL(k_pcg_p, dim3(ntile_total), dim3(kRedBlock), 0, s, d_sys,
                     nsys, it, part, state, sc);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

// ---------------------------------------------------------------------
// Batched Newton bookkeeping (one tile = kTileRows equations and
// kTileRows constrained dofs of one problem)
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(kRedBlock)
    k_residual_batch(const ResSys* __restrict__ sys, int nsys,
                     double* __restrict__ part) {
  __shared__ double s_red[kRedBlock];
  const int si = owner(sys, nsys, blockIdx.x);
  const ResSys& S = sys[si];
  const int t = blockIdx.x - S.tile0;
  const int k = t * kTileRows + threadIdx.x;
  double res = 0.0, rea = 0.0, load = 0.0;
  if (threadIdx.x < kTileRows) {
    if (k < S.neq) {
      const double v = -S.ifv[S.eqdof[k]];
      S.rhs[k] = v;
      res = v * v;
    }
    if (k < S.ncdof) {
      const double v = S.ifv[S.cdof[k]];
      rea = v * v;
    }
    if (k < S.nldof) load = S.ifv[S.ldof[k]];
  }
  res = block_sum(res, s_red);
  rea = block_sum(rea, s_red);
  load = block_sum(load, s_red);
  if (threadIdx.x == 0) {
    double* P = part + 3 * (int64_t)S.tile0;
    P[t] = res;
    P[S.ntile + t] = rea;
    P[2 * S.ntile + t] = load;
  }
}

__global__ void __launch_bounds__(kRedBlock)
    k_residual_sums(const ResSys* __restrict__ sys,
                    const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s_red[kRedBlock];
  const ResSys& S = sys[blockIdx.x];
  const double* P = part + 3 * (int64_t)S.tile0;
  const double res = sum_partials(P, S.ntile, s_red);
  const double rea = sum_partials(P + S.ntile, S.ntile, s_red);
  const double load = sum_partials(P + 2 * S.ntile, S.ntile, s_red);
  if (threadIdx.x == 0) {
    double* o = out + 8 * (int64_t)blockIdx.x;
    o[0] = res;
    o[1] = rea;
    o[2] = load;
    for (int f = 0; f < 4; ++f) o[3 + f] = (double)S.flags[f];
    o[7] = 0.0;
  }
}

__global__ void __launch_bounds__(kRedBlock)
    k_du_batch(const ResSys* __restrict__ sys, int nsys) {
  const int si = owner(sys, nsys, blockIdx.x);
  const ResSys& S = sys[si];
  if (!S.solve) return;
  const int t = blockIdx.x - S.tile0;
  const int k = t * kTileRows + threadIdx.x;
  if (threadIdx.x < kTileRows && k < S.neq) {
    const int d = S.eqdof[k];
    S.du[d] = S.du[d] + S.x[k];
  }
}

__global__ void k_axpy1(int64_t n, double* __restrict__ u,
                        const double* __restrict__ du) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[i] = u[i] + du[i];
}

__global__ void k_start_increment(const int32_t* __restrict__ eqdof, int neq,
                                  const int32_t* __restrict__ cdof,
                                  const double* __restrict__ cimp, int ncdof,
                                  double sf, double disp,
                                  double* __restrict__ du) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < neq) {
    const int d = eqdof[k];
    du[d] = sf * du[d];
  }
  if (k < ncdof) du[cdof[k]] = disp * cimp[k];
}

int launch_start_increment(const int32_t* eqdof, int neq, const int32_t* cdof,
                           const double* cimp, int ncdof, double sf,
                           double disp, double* du, hipStream_t s) {
  const int n = neq > ncdof ? neq : ncdof;
  if (n > 0)
    hipLaunchKernelGGL(k_start_increment, dim3((n + 255) / 256), dim3(256), 0,
                       s, eqdof, neq, cdof, cimp, ncdof, sf, disp, du);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_residual_batch(const ResSys* d_sys, int nsys, int ntile_total,
                          double* part, double* out, hipStream_t s) {
  if (nsys <= 0) return W3D_OK;
  if (ntile_total > 0)
    hipLaunchKernelGGL(k_residual_batch, dim3(ntile_total), dim3(kRedBlock),
                       0, s, d_sys, nsys, part);
  hipLaunchKernelGGL(k_residual_sums, dim3(nsys), dim3(kRedBlock), 0, s,
                     d_sys, part, out);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_du_batch(const ResSys* d_sys, int nsys, int ntile_total,
                    hipStream_t s) {
  if (nsys <= 0 || ntile_total <= 0) return W3D_OK;
  hipLaunchKernelGGL(k_du_batch, dim3(ntile_total), dim3(kRedBlock), 0, s,
                     d_sys, nsys);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int launch_axpy1(int64_t n, double* u, const double* du, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_axpy1, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, s, n, u, du);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

namespace {
inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
}

int solver_csr_values(const double* diag, const double* coef, int neq,
                      const uint32_t* vidx, int64_t nnz, double* vals,
                      hipStream_t s) {
  if (nnz > 0)
    hipLaunchKernelGGL(k_csr_values, dim3(nblk(nnz, 256)), dim3(256), 0, s,
                       diag, coef, neq, vidx, nnz, vals);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int solver_spmv(int n, const int64_t* rp, const int32_t* col,
                const double* val, const double* x, double* y, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_spmv, dim3(nblk((int64_t)n * 16, 256)), dim3(256), 0,
                       s, n, rp, col, val, x, y);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

// out[0] = a . b, with part of >= solver_npart(n) doubles of scratch
int solver_npart(int n) {
  int np = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
  return np < 1 ? 1 : (np > 1024 ? 1024 : np);
}
int solver_dot(int n, const double* a, const double* b, double* part,
               double* out, hipStream_t s) {
  int np = solver_npart(n);
  hipLaunchKernelGGL(k_dot_partial, dim3(np), dim3(kRedBlock), 0, s, n, a, b,
                     part);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kRedBlock), 0, s, np,
                     part, out);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int solver_jacobi(int n, const double* diag, const double* r, double* z,
                  hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_jacobi, dim3(nblk(n, 256)), dim3(256), 0, s, n, diag,
                       r, z);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int solver_update_p(int n, const double* z, const double* sc, double* p,
                    hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_update_p, dim3(nblk(n, 256)), dim3(256), 0, s, n, z,
                       sc, p);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int solver_update_xr(int n, const double* p, const double* q, const double* sc,
                     double* x, double* r, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_update_xr, dim3(nblk(n, 256)), dim3(256), 0, s, n, p,
                       q, sc, x, r);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int solver_residual(int n, const double* b, double* r, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_residual, dim3(nblk(n, 256)), dim3(256), 0, s, n, b,
                       r);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

int solver_shift(double* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_shift, dim3(1), dim3(1), 0, s, sc);
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

}  // namespace w3d
