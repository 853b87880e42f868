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
//     ranked row-major order described at ek_pos (base + i*utsz + word);
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
// words per recovery chunk and per block of the energy sums: internal
// energy, plastic work, sum of |element force| (addifv's sum_ifv)
constexpr int kBlkSums = 3;

// One element block as the kernels see it.
struct BlockDesc {
  int32_t span, felem, etype, order, ngp, nnode, totdof, utsz;
  int32_t mat, model, hist, bbar;
  int32_t table;     // index of the Gauss/derivative table
  int32_t ncry;      // mm10: crystals per point
  int32_t geonl;     // finite strains and rotations (lprops(18))
  int32_t trn;       // some node of the block has a skewed constraint
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

// Thermal expansion and the mm01 segmental curve set of a material
// (rknstr.f:150,1046-1053; set_up_segmental / set_up_h_prime).  curve_type
// 0 none, 1 one temperature-independent curve, 2 temperature dependent.
constexpr int kMaxCurves = 8;
struct ThermDesc {
  double alpha[6], alpha_zero;
  int32_t curve_type, ncurves;
  double curve_eps[2];
  double temp[kMaxCurves], e[kMaxCurves], nu[kMaxCurves],
      alpha_c[kMaxCurves], sig[kMaxCurves][2];
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
  double sf[kMaxGp][kMaxNodes];     // shape functions (temperatures)
  // slopes of the least-squares linear fit c1 xi + c2 eta + c3 zeta + c4
  // of values at the points: c_d = sum_p lsq[d][p] v_p (d = 0..2)
  double lsq[3][kMaxGp];
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
  double* rot_tmp;        // mm10 finite strains: [R n+1] of this recovery
                          // (9, points) per block, column-major
  double* jac_tmp;        // mm10 finite strains: each element's n = 0
                          // Jacobian at its last point (9, elements)
  double eps_bbar;
  double dt;              // time increment of the load step
  // finite strains: [R n+1] of every point (9 words, column-major, SoA
  // like the stresses: word k of point q of block b at
  // 9 q_off + k S G + q), and the displacements of the last recovery,
  // which place the tangent's configuration at X + u + du
  double* rot_n1;
  const double* u_t;
  const double* du_t;
  int32_t asym;           // asymmetric assembly: mm10 tangent unsymmetrised
  // skewed constraints: per node the index of its transformation (-1:
  // none) and the matrices, mat(r, c) of index t at trn_mat[9 t + r + 3 c]
  // (u_c = mat u_g)
  const int32_t* trn_idx;
  const double* trn_mat;
  // temperatures of the load step: temper_nodes, dtemp_nodes (nonode
  // each), temper_elems, dtemp_elems (noelem each), back to back; NULL
  // when none were set (all zero)
  const double* temps;
  const ThermDesc* therm;  // per material
  int32_t nonode, noelem;
  int32_t temps_incr;      // the step imposes temperature changes
  int32_t curves;          // some material has a segmental curve set
};

// Rotations of one node's 3-vector by its skewed-constraint matrix t
// (column-major): trnvec's switch 1 (v = T v, global to constraint
// compatible) and switch 2 (v = T^T v).
__host__ __device__ inline void trn_apply(const double* t, double* v) {
  const double x = v[0], y = v[1], z = v[2];
  v[0] = t[0] * x + t[3] * y + t[6] * z;
  v[1] = t[1] * x + t[4] * y + t[7] * z;
  v[2] = t[2] * x + t[5] * y + t[8] * z;
}
__host__ __device__ inline void trn_apply_t(const double* t, double* v) {
  const double x = v[0], y = v[1], z = v[2];
  v[0] = t[0] * x + t[1] * y + t[2] * z;
  v[1] = t[3] * x + t[4] * y + t[5] * z;
  v[2] = t[6] * x + t[7] * y + t[8] * z;
}

// Kernel launchers (w3d_kernels.hip, w3d_tangent.hip).
// Recovery passes: pass 0 runs mm01 blocks in full and the strain
// increments of mm10 blocks; pass 1 forms the element forces of mm10
// blocks from the stresses their material update left.
int launch_recover(const DevModel& M, const std::vector<BlockDesc>& hblocks,
                   const WorkItem* d_work, const std::vector<int>& groups,
                   const std::vector<int>& group_start,
                   const std::vector<int>& group_count, int step, int iter,
                   int nonlinear, const double* d_u, const double* d_du,
                   int want_energy, double* en_wg, int pass, hipStream_t s);
// mm10 update, one thread per point over work items of <= 64 points,
// persistent: grid workgroups take items from the counter d_counter.
int launch_mm10(const DevModel& M, const WorkItem* d_work, int nwork,
                int* d_counter, int grid, int step, int iter0_off,
                hipStream_t s);
// n int32 words of d set to v (a kernel, so that it may be captured in a
// graph in place of a memset)
int launch_fill_i32(int32_t* d, int n, int32_t v, hipStream_t s);
// grad Fe words of the small-strain mm10 blocks listed in d_blks whose
// index lies in [lo, hi) (rknstr_finish_cp), one workgroup per block.
int launch_mm10_grads(const DevModel& M, const int32_t* d_blks, int nblks,
                      int lo, int hi, int step, hipStream_t s);
int launch_ifv_gather(const double* einfvec, const int64_t* inv_ptr,
                      const int64_t* inv_off, int nonode, double* ifv,
                      hipStream_t s);
int launch_energy_reduce(const double* en_wg, const int32_t* blk_wfirst,
                         const int32_t* blk_wcount, int nblk,
                         double* blk_sums, hipStream_t s);
// Assembly over the segment map built by w3d_pattern: per segment a
// header (contributions | component mask << 8), its contributions' first
// Ke words in src, and per 64 segments the index of the first contribution
// and of the first output (diagonal outputs first, then coefficients).
int launch_assemble(const double* ek64, const float* ek32, int fp32,
                    const uint16_t* hdr, const uint32_t* src,
                    const int64_t* wbase, int64_t nseg, int64_t neq,
                    double* diag, double* coef, hipStream_t s);
// Asymmetric assembly (w3d_asym.hip): full FP64 Ke of every element of
// the work list into ekf (block b's elements from ekf_off[b], T*T words
// each), and its gather into the values of the full CSR rows.
int launch_tangent_full(const DevModel& M, const WorkItem* d_work,
                        const std::vector<int>& groups,
                        const std::vector<int>& gstart,
                        const std::vector<int>& gcount,
                        const int64_t* d_ekf_off, double* d_ekf,
                        hipStream_t s);
constexpr int kAsmMaxRow = 768;  // longest CSR row the gather takes
// Jacobi-BiCGStab on full CSR rows from x = 0; work holds
// bicgstab_words(n) = 9 n + kBicgScalars + solver_npart(n) doubles.
// W3D_ERR_NUMERIC after maxit iterations or on a breakdown (a zero rho or
// (rh, v)), x then the last iterate.  Clears its work array with kernels
// (no memsets).
constexpr int kBicgScalars = 16;
int bicgstab_words(int n);
int bicgstab(int n, const int64_t* rp, const int32_t* col, const double* val,
             const double* b, double* x, double* work, double rtol, int maxit,
             int* iters, double* relres, hipStream_t st);
constexpr int kAsmNone = 0xFFFF;  // constrained column in the gather map
int launch_assemble_full(const double* ekf, const int64_t* ebase,
                         const int64_t* atoff, const uint16_t* at,
                         const int32_t* eq_dof, const int64_t* inv_ptr,
                         const int32_t* inv_ln, const int64_t* rp,
                         double* val, int neq, hipStream_t s);
int launch_tangent_mfma(const DevModel& M, const WorkItem* d_work,
                        const std::vector<int>& groups,
                        const std::vector<int>& group_start,
                        const std::vector<int>& group_count, int precision,
                        hipStream_t s);
// MFMA tangent: waves per workgroup and elements per workgroup (each wave
// takes batches of tan_batch consecutive elements)
__host__ __device__ constexpr int tan_mfma_waves(int nnode) {
  return (void)nnode, 4;
}
__host__ __device__ constexpr int tan_mfma_chunk(int nnode) {
  return nnode == 20 ? 16 : 32;
}
// Elements of a tangent wave's batch: as many as fill the 64 lanes with
// (element, point) pairs, at most 16; for the FP64 contraction at most 4,
// and one for rules with more than 8 points (a larger batch's LDS would
// cut the waves per SIMD).
__host__ __device__ constexpr int tan_batch(int ngp, int tbytes) {
  return tbytes == 8
             ? (ngp > 8 ? 1 : (64 / ngp > 4 ? 4 : 64 / ngp))
             : (64 / ngp < 1 ? 1 : (64 / ngp > 16 ? 16 : 64 / ngp));
}

// Jacobi-PCG building blocks (w3d_solver.hip)
int solver_csr_values(const double* diag, const double* coef, int neq,
                      const uint32_t* vidx, int64_t nnz, double* vals,
                      hipStream_t s);
int solver_spmv(int n, const int64_t* rp, const int32_t* col,
                const double* val, const double* x, double* y, hipStream_t s);
int solver_npart(int n);
// w3d_solve_pcg_static's work memory: a file-scope device array of
// 4 kStaticMaxEq + 8 + 1024 doubles, its address on `device` (the caller
// has made that device current)
constexpr int kStaticMaxEq = 1 << 19;
constexpr int kMaxDevices = 64;
int solver_static_scratch(int device, double** out);
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

// ---------------------------------------------------------------------
// Batches of independent problems (w3d_serve.cpp): one launch covers the
// work of every problem in the batch.
// ---------------------------------------------------------------------
// One problem's share of a batched mm10 update: its model, its work list
// and the predicates of its step and iteration.  Global item g of the
// launch belongs to the last slot with item0 <= g.
struct Mm10Slot {
  DevModel M;
  const WorkItem* work;
  int32_t item0;
  int32_t nwork;
  int32_t step;
  int32_t iter0_off;
};
int launch_mm10_batch(const Mm10Slot* d_slots, int nslot, int nwork_total,
                      int* d_counter, int grid, hipStream_t s);

// Rows of one tile of a batched vector kernel (one workgroup each).
constexpr int kTileRows = 128;

// One system of a batched Jacobi-PCG: the full CSR of its assembled
// matrix, its diagonal, right-hand side and vectors, and its tiles.
// Per system the batch keeps 7 partial-sum rows of ntile words (b.b, then
// p.q, r.r and r.z by iteration parity), the state (done: 0 running,
// 1 converged, 2 iteration limit; iterations) and the scalars (|b|, r.r).
struct PcgSys {
  const int64_t* rp;
  const int32_t* col;
  const double* val;
  const double* diag;
  const double* b;
  double* x;
  double* r;
  double* p;
  double* q;
  double rtol;            // |r| <= rtol |b| ends the solve
  int32_t maxit;          // or this many iterations
  int32_t n;
  int32_t tile0;
  int32_t ntile;
};
// x = 0, r = b, p = D^-1 b and the partial sums; then |b| per system
int launch_pcg_batch_init(const PcgSys* d_sys, int nsys, int ntile_total,
                          double* part, int32_t* state, double* sc,
                          hipStream_t s);
// iteration it (0-based) of every running system: q = A p; x, r; p
int launch_pcg_batch_iter(const PcgSys* d_sys, int nsys, int ntile_total,
                          int it, double* part, int32_t* state, double* sc,
                          hipStream_t s);

// One problem's Newton bookkeeping for the batched residual and update:
// rhs = -ifv on the equations, |rhs|^2, the reaction |ifv|^2 on the
// constrained dofs, and the sum of ifv over the loaded dofs; du += x on
// the equations of the solved systems.
struct ResSys {
  const double* ifv;
  const int32_t* eqdof;   // dof of each equation
  const int32_t* cdof;    // constrained dofs
  const int32_t* ldof;    // loaded dofs (a subset of cdof)
  const int32_t* flags;   // the context's recovery flags
  double* rhs;
  double* du;
  const double* x;
  int32_t neq, ncdof, nldof;
  int32_t tile0, ntile;
  int32_t solve;          // du += x for this system (launch_du_batch)
};
// out[8 s + k]: k = 0 |rhs|^2, 1 reaction^2, 2 load, 3.. 6 the flags
int launch_residual_batch(const ResSys* d_sys, int nsys, int ntile_total,
                          double* part, double* out, hipStream_t s);
int launch_du_batch(const ResSys* d_sys, int nsys, int ntile_total,
                    hipStream_t s);
// u += du (n words)
int launch_axpy1(int64_t n, double* u, const double* du, hipStream_t s);
// the start of a load increment: du *= sf on the equations' dofs, and
// du = disp * cimp on the constrained dofs (cimp: the full step's imposed
// increments, in cdof order)
int launch_start_increment(const int32_t* eqdof, int neq, const int32_t* cdof,
                           const double* cimp, int ncdof, double sf,
                           double disp, double* du, hipStream_t s);

// Host set-up of one crystal record from the material and its angles
// (w3d_mm10_setup.cpp); false for an unsupported slip / elastic type.
bool mm10_crystal_setup(const w3d_material& m, const double* angles,
                        CrystalDev* cr);

// Kernel-group keys: one instantiation per (nnode, ngp, bbar), per
// material family for the recovery (mm10 blocks run in two passes), per
// kinematics (finite strains) and, for the tangent, by whether a node has
// a skewed constraint.  Bit 0: geonl; bit 1: mm10; bit 2: skewed
// constraints; key >> 3 the element (nodes, points, B-bar).
int group_key(int nnode, int ngp, int bbar, int mm10 = 0, int geonl = 0,
              int trn = 0);

// Lower-triangle index of (row, col) in the 21-word [D] storage
// (rstgp1_store_cep order (1,1),(2,1),(2,2),(3,1),...).
__host__ __device__ constexpr int tri(int r, int c) {
  return r >= c ? r * (r + 1) / 2 + c : c * (c + 1) / 2 + r;
}

// Element tangents are stored in a ranked row-major layout: the element's
// nodes are ranked by ascending global number (ties by local index), its
// dofs renumbered d' = 3*rank + component, and the upper triangle of Ke in
// that numbering stored row by row: word(p', q') = p' T - p'(p'-1)/2 +
// q' - p' for p' <= q' (T = 3 NN).  The three rows of the node of rank r
// are then the contiguous run of words the element adds to that node's
// rows of the assembled upper triangle, each row ordered like the global
// columns, so the row-by-row assembly reads them in runs.
__host__ __device__ constexpr int ek_word(int T, int p, int q) {
  return p * T - p * (p - 1) / 2 + q - p;
}
// Rank of every local node: the number of local nodes with a smaller
// global number, or an equal one at a smaller local index.
__host__ __device__ inline void node_ranks(const int32_t* nodes, int nn,
                                           int* rank) {
  for (int a = 0; a < nn; ++a) {
    int r = 0;
    for (int b = 0; b < nn; ++b)
      r += (nodes[b] < nodes[a] || (nodes[b] == nodes[a] && b < a)) ? 1 : 0;
    rank[a] = r;
  }
}
// Word of Ke(p, q) (local dofs, dof = component*NN + node) given the ranks.
__host__ __device__ inline int ek_pos(int NN, const int* rank, int p, int q) {
  int a = 3 * rank[p % NN] + p / NN, b = 3 * rank[q % NN] + q / NN;
  return a <= b ? ek_word(3 * NN, a, b) : ek_word(3 * NN, b, a);
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

// MI355X dispatches workgroup b of a launch to XCD b % 8, and each XCD has
// its own L2.  For a kernel whose neighbouring workgroups read the same
// lines (the force gather's element vectors), workgroup b takes the
// logical index that makes every XCD run one contiguous range of the
// work: XCD x gets the x-th of eight nearly equal ranges, in order.  A
// bijection on [0, nb) for any nb.
constexpr int kNumXcd = 8;
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
  const int64_t x = b % kNumXcd, i = b / kNumXcd;
  const int64_t q = nb / kNumXcd, r = nb % kNumXcd;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}


}  // namespace w3d
