// w3d_internal.h -- shared declarations of libw3dblk (host and device).
//
// Device data layout (HBM), per element block b of span S and ngp G:
//   * Gauss-point state is structure-of-arrays inside the block: word k of
//     point (element i, point g) sits at base_b + k*S*G + i*G + g, so the
//     lanes of a wavefront that walk consecutive (i, g) read and write
//     consecutive addresses.
//   * element forces einfvec are element-contiguous and node-major inside
//     the element: base + i*T + 3*j + c (component c of local node j).
//   * element tangents are element-contiguous, utsz words each, in the
//     node-pair block order described at npb_pos (base + i*utsz + word);
//     w3d_get_block_state converts them to the reference's packed upper
//     triangle (estiff_blocks (utsz, span)).
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
// mm10 history of one point: 87 common words + 71 per crystal (mm10_d.f)
constexpr int kMM10Common = 87;
constexpr int kMM10Crystal = 71;

// One element block as the kernels see it.
struct BlockDesc {
  int32_t span, felem, etype, order, ngp, nnode, totdof, utsz;
  int32_t mat, model, hist, bbar;
  int32_t table;     // index of the Gauss/derivative table
  int32_t ncry;      // mm10: crystals per point
  int64_t q_off;     // first Gauss point of the block (global numbering)
  int64_t hist_off;  // first history word of the block
  int64_t ek_off;    // first Ke word of the block
  int64_t ifv_off;   // first einfvec word of the block
  int64_t conn_off;  // first connectivity entry of the block
  int64_t crys_off;  // mm10: first crystal record of the block
};

// mm10 (Voce) constants of a material (the c_array entry).
struct CpDesc {
  double rate_n, tau_y, tau_v, voce_m, theta_0, iD_v, eps_dot_0_y;
  double atol, atol1, rtol, rtol1;
  int32_t miter, alter_mode, angle_type, ncry;
};

// Material constants; mm01's widened the way setup_mm01_rknstr does
// (REAL -> FP64).
struct MatDesc {
  int32_t model;
  int32_t lnelas;
  double e, nu, beta, hprime, yld;
  CpDesc cp;
};

// One crystal of one element (setup_mm10_rknstr): stiffness rotated into
// the reference frame ((i,j) at C[i + 6j]), Schmid tensors (Voigt,
// engineering shear) and skew vectors of the 12 slip systems, the crystal
// -> reference rotation g ((i,j) at g[i + 3j]) and the initial angles.
struct CrystalDev {
  double C[36];
  double ms[12][6];
  double qs[12][3];
  double g[9];
  double angles[3];
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
  const CrystalDev* crys; // mm10 crystal records
  double* uddt;           // mm10: strain increments (6, points) per block
  double eps_bbar;
  double dt;              // time increment of the load step
  // word -> (rank r, rank s, component i, component j) of the node-pair
  // block layout, r | s << 5 | i << 10 | j << 12, for 8, 10 and 20 nodes
  // at npb_tab_off(NN)
  const uint16_t* npb_tab;
};

__host__ __device__ constexpr int npb_tab_off(int NN) {
  return NN == 8 ? 0 : NN == 10 ? 300 : 765;
}
constexpr int kNpbTabSize = 300 + 465 + 1830;

// Kernel launchers (w3d_kernels.hip, w3d_tangent.hip).
// Recovery passes: pass 0 runs mm01 blocks in full and the strain
// increments of mm10 blocks; pass 1 forms the element forces of mm10
// blocks from the stresses their material update left.
int launch_recover(const DevModel& M, const std::vector<BlockDesc>& hblocks,
                   const WorkItem* d_work, const std::vector<int>& groups,
                   const std::vector<int>& group_start,
                   const std::vector<int>& group_count, int step, int iter,
                   int nonlinear, const double* d_du, int want_energy,
                   double* en_wg, int pass, hipStream_t s);
// mm10 update, one thread per point over work items of <= 64 points.
int launch_mm10(const DevModel& M, const WorkItem* d_work, int nwork,
                int step, int iter0_off, hipStream_t s);
int launch_ifv_gather(const double* einfvec, const int64_t* inv_ptr,
                      const int64_t* inv_off, int nonode, double* ifv,
                      hipStream_t s);
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

// Jacobi-PCG building blocks (w3d_solver.hip)
int solver_csr_values(const double* diag, const double* coef, int neq,
                      const uint32_t* vidx, int64_t nnz, double* vals,
                      hipStream_t s);
int solver_spmv(int n, const int64_t* rp, const int32_t* col,
                const double* val, const double* x, double* y, hipStream_t s);
int solver_npart(int n);
int solver_dot(int n, const double* a, const double* b, double* part,
               double* out, hipStream_t s);
int solver_jacobi(int n, const double* diag, const double* r, double* z,
                  hipStream_t s);
int solver_update_p(int n, const double* z, const double* sc, double* p,
                    hipStream_t s);
int solver_update_xr(int n, const double* p, const double* q,
                     const double* sc, double* x, double* r, hipStream_t s);
int solver_residual(int n, const double* b, double* r, hipStream_t s);
int solver_shift(double* sc, hipStream_t s);

// Host set-up of one crystal record from the material and its angles
// (w3d_mm10_setup.cpp); false for an unsupported slip / elastic type.
bool mm10_crystal_setup(const w3d_material& m, const double* angles,
                        CrystalDev* cr);

// Kernel-group keys: one instantiation per (nnode, ngp, bbar) and, for the
// recovery, per material family (mm10 blocks run in two passes).
int group_key(int nnode, int ngp, int bbar, int mm10 = 0);

// Lower-triangle index of (row, col) in the 21-word [D] storage
// (rstgp1_store_cep order (1,1),(2,1),(2,2),(3,1),...).
__host__ __device__ constexpr int tri(int r, int c) {
  return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r;
}

// Element tangents are stored in a node-pair block layout: the element's
// nodes ranked by ascending global number; block row r (the node of rank r)
// holds the upper triangle of its 3x3 diagonal block (6 words, row-major:
// (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)) then the full 3x3 blocks with every
// node of higher rank s = r+1 .. NN-1 (9 words each, row-major, rows the
// components of the rank-r node).  The words of one element are exactly the
// contributions it makes to the rows of each of its nodes in the assembled
// upper triangle, so the row-by-row assembly reads them in runs.  Same size
// as the packed upper triangle: 6 NN + 9 NN (NN-1)/2 = 3NN (3NN+1)/2.
__host__ __device__ constexpr int npb_row_start(int NN, int r) {
  return 6 * r + 9 * (r * (NN - 1) - r * (r - 1) / 2);
}
// Word of Ke(p, q) (local dofs, dof = component*NN + node) given the ranks.
__host__ __device__ inline int npb_pos(int NN, const int* rank, int p, int q) {
  int a = p % NN, i = p / NN, b = q % NN, j = q / NN;
  int ra = rank[a], rb = rank[b];
  if (ra == rb) {
    int lo = i < j ? i : j, hi = i < j ? j : i;
    return npb_row_start(NN, ra) + lo * (5 - lo) / 2 + hi;
  }
  if (ra > rb) {
    int t = ra; ra = rb; rb = t;
    t = i; i = j; j = t;
  }
  return npb_row_start(NN, ra) + 6 + 9 * (rb - ra - 1) + 3 * i + j;
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
