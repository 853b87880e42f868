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

#include "w3d_internal.h"

namespace w3d {

constexpr int kRedBlock = 256;

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
