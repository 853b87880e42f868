// w3d_internal.h -- shared declarations of libw3dblk (host and device).
//
// Device data layout (HBM), per element block b of span S and ngp G:
//   * Gauss-point state is structure-of-arrays inside the block: word k of
//     point (element i, point g) sits at base_b + k*S*G + i*G + g, so the
//     lanes of a wavefront that walk consecutive (i, g) read and write
//     consecutive addresses.
//   * element forces einfvec are element-contiguous: base + i*T + dof.
//   * element tangents are element-contiguous packed upper triangles, the
//     reference's estiff_blocks (utsz, span) layout: base + i*utsz + pos.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/w3d.h"

namespace w3d {

constexpr int kMaxNodes = 20;
constexpr int kMaxDof = 60;
constexpr int kMaxGp = 27;
constexpr int kMM01Hist = 11;
constexpr int kMM10Hist = 158;

// One element block as the kernels see it.
struct BlockDesc {
  int32_t span, felem, etype, order, ngp, nnode, totdof, utsz;
  int32_t mat, model, hist, bbar;
  int32_t table;     // index of the Gauss/derivative table
  int32_t pad_;
  int64_t q_off;     // first Gauss point of the block (global numbering)
  int64_t hist_off;  // first history word of the block
  int64_t ek_off;    // first Ke word of the block
  int64_t ifv_off;   // first einfvec word of the block
  int64_t conn_off;  // first connectivity entry of the block
};

// Material constants widened the way setup_mm01_rknstr does (REAL -> FP64).
struct MatDesc {
  int32_t model;
  int32_t lnelas;
  double e, nu, beta, hprime, yld;
  double mm10[24];
};

// Gauss points and parametric derivatives of one (element type, order).
struct GaussTable {
  int32_t etype, order, ngp, nnode;
  double w[kMaxGp];
  double dn[3][kMaxGp][kMaxNodes];  // d/dxi, d/deta, d/dzeta
};

// A chunk of consecutive elements of one block handled by one workgroup.
struct WorkItem {
  int32_t blk;
  int32_t first;  // 0-based element inside the block
  int32_t count;
  int32_t pad_;
};

// Pointers handed to the kernels (all device memory).
struct DevModel {
  const BlockDesc* blocks;
  const MatDesc* mats;
  const GaussTable* tables;
  const double* coords;   // (3*nonode)
  const int32_t* conn;    // 0-based node of (block, element, local node)
  double* stress_n;
  double* stress_n1;
  double* strain_n;
  double* strain_n1;
  double* hist_n;
  double* hist_n1;
  double* cep;
  double* ek64;
  float* ek32;
  double* einfvec;
  double* evol;
  int32_t* flags;         // [0] fatal numeric, [1] cut step, [2] det J <= 0
  double eps_bbar;
};

// Kernel launchers (w3d_kernels.hip, w3d_tangent.hip).
int launch_recover(const DevModel& M, const std::vector<BlockDesc>& hblocks,
                   const WorkItem* d_work, const std::vector<int>& groups,
                   const std::vector<int>& group_start,
                   const std::vector<int>& group_count, int step, int iter,
                   int nonlinear, const double* d_du, int want_energy,
                   double* en_wg, hipStream_t s);
int launch_ifv_gather(const double* einfvec, const int64_t* inv_ptr,
                      const int64_t* inv_off, const int32_t* inv_nn,
                      int nonode, double* ifv, hipStream_t s);
int launch_energy_reduce(const double* en_wg, const int32_t* blk_wfirst,
                         const int32_t* blk_wcount, int nblk,
                         double* blk_sums, hipStream_t s);
int launch_assemble(const double* ek64, const float* ek32, int fp32,
                    const int64_t* row_ptr, const uint32_t* src,
                    int64_t nout, int64_t neq, double* diag, double* coef,
                    hipStream_t s);
int launch_tangent_mfma(const DevModel& M, const WorkItem* d_work,
                        const std::vector<int>& groups,
                        const std::vector<int>& group_start,
                        const std::vector<int>& group_count, int precision,
                        hipStream_t s);
// elements per workgroup of the MFMA tangent (one wave per element)
__host__ __device__ constexpr int tan_mfma_waves(int nnode) {
  return nnode == 20 ? 2 : 4;
}
__host__ __device__ constexpr int tan_mfma_chunk(int nnode) {
  return 8 * tan_mfma_waves(nnode);
}

// Kernel-group keys: one instantiation per (nnode, ngp, bbar).
int group_key(int nnode, int ngp, int bbar);

// Lower-triangle index of (row, col) in the 21-word [D] storage
// (rstgp1_store_cep order (1,1),(2,1),(2,2),(3,1),...).
__host__ __device__ constexpr int tri(int r, int c) {
  return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r;
}

// Upper 16x16 tiles (I <= J) of a tile grid, enumerated column by column.
__host__ __device__ constexpr int tile_i(int t) {
  return t < 1 ? 0 : t < 3 ? t - 1 : t < 6 ? t - 3 : t - 6;
}
__host__ __device__ constexpr int tile_j(int t) {
  return t < 1 ? 0 : t < 3 ? 1 : t < 6 ? 2 : 3;
}

// Elements per workgroup of the recovery (lanes = element x point).
__host__ __device__ constexpr int rec_elems(int ngp) {
  return 256 / ngp < 40 ? 256 / ngp : 40;
}


}  // namespace w3d
