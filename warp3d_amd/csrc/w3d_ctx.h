// w3d_ctx.h -- the library context (the opaque w3d_ctx of w3d.h) and the
// host helpers shared by the C-ABI translation units (w3d_api.cpp,
// w3d_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "w3d_internal.h"

namespace w3d {

// records msg as the calling thread's last error (w3d_last_error) and
// returns code
int fail(int code, const std::string& msg);

// bytes dalloc has allocated on the calling thread (set-up entry points
// charge the difference to their context)
extern thread_local int64_t t_dalloc_bytes;

template <typename T>
int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess)
    return fail(W3D_ERR_NOMEM, std::string("hipMalloc: ") +
                                   hipGetErrorString(e));
  t_dalloc_bytes += (int64_t)(n * sizeof(T));
  return W3D_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

template <typename T>
int upload(T** d, const std::vector<T>& h) {
  int rc = dalloc(d, h.size());
  if (rc) return rc;
  if (!h.empty()) {
    hipError_t e = hipMemcpy(*d, h.data(), h.size() * sizeof(T),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return fail(W3D_ERR_DEVICE, std::string("hipMemcpy: ") +
                                      hipGetErrorString(e));
  }
  return W3D_OK;
}

}  // namespace w3d

#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess)                                                \
      return w3d::fail(W3D_ERR_DEVICE, std::string(#x) + ": " +          \
                                           hipGetErrorString(e_));       \
  } while (0)

using w3d::BlockDesc;
using w3d::CrystalDev;
using w3d::DevModel;
using w3d::GaussTable;
using w3d::MatDesc;
using w3d::WorkItem;

struct w3d_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nonode = 0, noelem = 0, nelblk = 0;
  std::vector<int32_t> etype, order, matnum, bbar, geonl;
  bool has_geonl = false;
  // finite strains: [R] at n and n+1 of every point (SoA like the
  // stresses), and the displacements of the last recovery (the tangent's
  // configuration)
  double *d_rot_n = nullptr, *d_rot_n1 = nullptr;
  double *d_u_t = nullptr, *d_du_t = nullptr;
  std::vector<int32_t> elblks;
  // skewed constraints: per node the index of its matrix (-1: none), and
  // the matrices (column-major 3x3)
  std::vector<int32_t> trn_idx;
  std::vector<double> trn_mat;
  int32_t* d_trn_idx = nullptr;
  double* d_trn_mat = nullptr;
  std::vector<BlockDesc> blocks;
  std::vector<MatDesc> mats;
  // thermal expansion and segmental curves per material; the step's
  // temperatures (nodes, node changes, elements, element changes) once set
  std::vector<w3d::ThermDesc> therm;
  w3d::ThermDesc* d_therm = nullptr;
  bool curves = false;
  double* d_temps = nullptr;
  int temps_incr = 0;
  std::vector<GaussTable> tables;
  std::vector<int32_t> conn;      // 0-based nodes, block order
  std::vector<int32_t> elem_blk;  // block of each element (0-based)
  // inverse incidences: per node, (element, local node) ascending element
  std::vector<int64_t> inv_ptr;
  std::vector<int32_t> inv_elem, inv_ln;
  int64_t nq = 0, nhist = 0, nek = 0, nifv = 0;
  // device memory
  BlockDesc* d_blocks = nullptr;
  MatDesc* d_mats = nullptr;
  GaussTable* d_tables = nullptr;
  double* d_coords = nullptr;
  int32_t* d_conn = nullptr;
  double *d_stress_n = nullptr, *d_stress_n1 = nullptr;
  double *d_strain_n = nullptr, *d_strain_n1 = nullptr;
  double *d_hist_n = nullptr, *d_hist_n1 = nullptr;
  double* d_cep = nullptr;
  double* d_ek64 = nullptr;
  float* d_ek32 = nullptr;
  double *d_einfvec = nullptr, *d_evol = nullptr, *d_en_wg = nullptr;
  int32_t* d_flags = nullptr;
  // mm10: crystal records, strain increments, per-point work items
  CrystalDev* d_crys = nullptr;
  double* d_uddt = nullptr;
  double* d_rot_tmp = nullptr;  // mm10 with finite strains only
  double* d_jac_tmp = nullptr;
  WorkItem* d_work_mm10 = nullptr;
  int n_work_mm10 = 0;
  int32_t* d_mm10_blks = nullptr;  // indices of the mm10 blocks
  int n_mm10_blks = 0;
  int32_t* d_counter = nullptr;     // work counter of persistent kernels
  int mm10_grid = 0;                // workgroups of the persistent mm10
  int n_cu = 0;                     // compute units of the device
  // tangent work-item size (elements) per kernel group, autotuned
  std::vector<std::pair<int, int>> tan_chunk;
  // captured Newton iterations (w3d_iteration_dev), newest last
  struct GraphEntry {
    int key[5];
    const void* ptr[5];
    double dt;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;
  bool capturing = false;
  int last_step = 1;                // step of the last recovery
  bool has_mm10 = false;
  double dt = 1.0;
  int64_t dev_bytes = 0;            // device bytes of set-up and pattern
  WorkItem *d_work_rec = nullptr, *d_work_tan = nullptr;
  std::vector<int> rec_groups, rec_start, rec_count;
  std::vector<int> tan_groups, tan_start, tan_count;
  int64_t *d_inv_ptr = nullptr, *d_inv_off = nullptr;
  int32_t *d_blk_wfirst = nullptr, *d_blk_wcount = nullptr;
  // per block: internal energy, plastic work and the sum of |element
  // force| of the last recovery (kBlkSums words)
  double* d_blk_sums = nullptr;
  // the last whole-pass recovery's addifv sums (sum_ifv, num_term_ifv)
  double sum_ifv = 0.0;
  int64_t num_term_ifv = 0;
  // per-block calls (w3d_recover_block, w3d_tangent_block), safe to run
  // concurrently for distinct blocks: each block's run of work items in
  // the recovery, tangent and mm10 lists, its own flags and mm10 counter,
  // and a pool of streams the calls borrow
  struct BlockRuns {
    int rec_key = 0, rec_first = 0, rec_n = 0;
    int tan_key = 0, tan_first = 0, tan_n = 0;
    int mm10_first = 0, mm10_n = 0, mm10_idx = -1;
  };
  std::vector<BlockRuns> runs;
  int32_t* d_blk_flags = nullptr;    // 4 words per block
  int32_t* d_blk_counter = nullptr;  // one word per block
  std::mutex pool_mu;
  std::vector<hipStream_t> pool_idle, pool_all;
  bool disp_set = false;             // w3d_set_displacements was called
  double *d_u = nullptr, *d_du = nullptr, *d_ifv = nullptr;
  // sparsity + assembly map
  int neqns = 0;
  int64_t ncoeff = 0;
  std::vector<int32_t> dof_eqn_map, eqn_node_map, k_ptrs, k_indexes;
  std::vector<int64_t> row_start;
  uint16_t* d_asm_hdr = nullptr;   // assembly segments (w3d_pattern)
  uint32_t* d_asm_src = nullptr;
  int64_t* d_asm_wbase = nullptr;
  int64_t asm_nseg = 0;
  bool asm_ready = false;
  // the symmetric assembly (assem_by_row) into (diag, coef) on stream s
  int enqueue_assemble(int fp32, double* diag, double* coef,
                       hipStream_t s) const {
    return w3d::launch_assemble(d_ek64, d_ek32, fp32, d_asm_hdr, d_asm_src,
                                d_asm_wbase, asm_nseg, neqns, diag, coef, s);
  }
  double* d_kout = nullptr;  // neqns diagonal then ncoeff coefficients
  int64_t assembled = 0;     // serial of the last assembly
  // full symmetric CSR of the assembled matrix for the solver: row
  // pointers, columns (0-based), the d_kout index of each entry, values
  int64_t csr_nnz = 0, csr_serial = -1;
  int64_t* d_csr_rp = nullptr;
  int32_t* d_csr_col = nullptr;
  uint32_t* d_csr_vidx = nullptr;
  double* d_csr_val = nullptr;
  double* d_pcg = nullptr;   // solver work vectors and scalars
  // BiCGStab of the asymmetric CSR: b, x and the work vectors, allocated
  // with the gather map, once per pattern
  double* d_bicg = nullptr;
  // asymmetric assembly (w3d_set_asymmetric): the full Ke of every element
  // (block b from d_ekf_off[b]), whether the last tangent formed it, and
  // the gather tables of the CSR assembly (rebuilt after w3d_pattern)
  bool asym = false, asym_tangent = false;
  double* d_ekf = nullptr;
  int64_t* d_ekf_off = nullptr;
  int64_t n_ekf = 0;
  int32_t *d_fa_eqdof = nullptr, *d_fa_ln = nullptr;
  int64_t *d_fa_invptr = nullptr, *d_fa_ebase = nullptr, *d_fa_atoff = nullptr;
  uint16_t* d_fa_at = nullptr;
  int csr_maxrow = 0;  // longest row of the full CSR pattern
  bool csr_asym_ready = false;  // d_csr_val holds the asymmetric CSR
  int tangent_precision = -1;
  double eps_bbar = 0.0;
  // phase timing (w3d_set_profiling)
  bool profiling = false;
  // up to two timed segments per phase between collections
  hipEvent_t ev_a[W3D_NUM_PHASES][2] = {}, ev_b[W3D_NUM_PHASES][2] = {};
  int nseg[W3D_NUM_PHASES] = {};
  int64_t prof_n[W3D_NUM_PHASES] = {};
  double prof_ms[W3D_NUM_PHASES] = {};

  DevModel dev() const {
    DevModel M;
    M.blocks = d_blocks;
    M.mats = d_mats;
    M.tables = d_tables;
    M.coords = d_coords;
    M.conn = d_conn;
    M.stress_n = d_stress_n;
    M.stress_n1 = d_stress_n1;
    M.strain_n = d_strain_n;
    M.strain_n1 = d_strain_n1;
    M.hist_n = d_hist_n;
    M.hist_n1 = d_hist_n1;
    M.cep = d_cep;
    M.ek64 = d_ek64;
    M.ek32 = d_ek32;
    M.einfvec = d_einfvec;
    M.evol = d_evol;
    M.flags = d_flags;
    M.crys = d_crys;
    M.uddt = d_uddt;
    M.rot_tmp = d_rot_tmp;
    M.jac_tmp = d_jac_tmp;
    M.eps_bbar = eps_bbar;
    M.dt = dt;
    M.rot_n1 = d_rot_n1;
    M.u_t = d_u_t;
    M.du_t = d_du_t;
    M.asym = asym ? 1 : 0;
    M.trn_idx = d_trn_idx;
    M.trn_mat = d_trn_mat;
    M.temps = d_temps;
    M.therm = d_therm;
    M.nonode = nonode;
    M.noelem = noelem;
    M.temps_incr = temps_incr;
    M.curves = curves ? 1 : 0;
    return M;
  }
};

namespace w3d {

// Work lists of one whole-pass recovery (the context's resident lists).
struct RecoverLists {
  const WorkItem* d_w = nullptr;
  const WorkItem* d_wm = nullptr;
  int nwm = 0;
  std::vector<int> sg, ss, sn;
};

int recover_lists(w3d_ctx* c, RecoverLists& L);
// Per-block calls: a stream of the context's pool, its return, the
// recovery of one block (0-based) enqueued on a stream, and the status of
// a block's four recovery flags.
int borrow_stream(w3d_ctx* c, hipStream_t* s);
void return_stream(w3d_ctx* c, hipStream_t s);
int enqueue_block_recover(w3d_ctx* c, int b, hipStream_t s, int step,
                          int iter, int extrapolated);
int flags_status(const int32_t* f, int step);
void build_runs(w3d_ctx* c);
// The stages of one recovery on the context's stream (no host
// synchronisation, no memsets): stage 0 clears the flags and runs the
// first recovery pass, stage 1 the mm10 update, stage 2 its grad Fe words,
// stage 3 the second pass, the structure-force gather into d_ifv (if
// given) and the per-block energy sums (if sums).  enqueue_recover runs
// all four.
// mm10's iteration-0 linear-elastic update (rstgp1.f:4453-4457)
inline int mm10_iter0_off(int step, int iter, int extrapolated) {
  return (iter == 0 && (!extrapolated || step == 1)) ? 1 : 0;
}
int enqueue_recover_stage(w3d_ctx* c, const RecoverLists& L, int stage,
                          int step, int iter, int extrapolated,
                          const double* d_u, const double* d_du, double* d_ifv,
                          bool sums, int first_blk, int nblk);
int enqueue_recover(w3d_ctx* c, const RecoverLists& L, int step, int iter,
                    int extrapolated, const double* d_u, const double* d_du,
                    double* d_ifv, bool sums, int first_blk, int nblk);
// The [n+1] -> [n] state copies of w3d_commit_step, and the [n] ->
// [n+1] copies of w3d_rollback, enqueued only.
int enqueue_commit(w3d_ctx* c);
int enqueue_rollback(w3d_ctx* c);
// CSR values of the solver from an assembly (diag, coef) on the stream.
int enqueue_csr_values(w3d_ctx* c, const double* d_diag, const double* d_coef);

}  // namespace w3d
