// w3d_serve.cpp -- batched serving of independent quasi-static problems
// (w3d_problem, w3d_batch of include/w3d.h).  No reference counterpart:
// the reference solves one model per process with a direct solver.
//
// One w3d_batch_advance is one Newton iteration of every running problem:
//   1. a problem whose step converged in the previous advance commits its
//      state (update.f:15 restated by enqueue_commit) and adds its increment
//      to u; the next step starts from the same increment (extrapolation);
//   2. recovery of every problem: the per-problem passes on the batch's
//      stream and one persistent mm10 launch over all the problems' points;
//   3. one residual launch over all problems (rhs = -ifv on the equations,
//      the norms), one copy to the host and the convergence decisions;
//   4. block tangent and assembly of the unconverged problems, then one
//      batched Jacobi-PCG over their systems, synchronising with the host
//      once per chunk of iterations;
//   5. one launch adds the corrections to du.
// A problem's kernels read only its own data and its reductions run over
// its own tiles in a fixed order, so the batch composition never changes
// its results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "w3d_ctx.h"

using namespace w3d;

// adapt_check.f's stack of sub-increments: rows (start, end, increment,
// extrapolation scale), the current one at row level - 1
struct AdaptStack {
  static constexpr int kDiv = 4, kRows = 64;
  double st[kRows][4] = {};
  int level = 1;
  double scaling = 1.0, disp = 1.0;
  // process type 3: a new step
  void init() {
    st[0][0] = 0.0;
    st[0][1] = 1.0;
    st[0][2] = 1.0;
    st[0][3] = 1.0;
    level = 1;
    scaling = 1.0;
    disp = 1.0;
  }
  // process type 1: split the current increment (false: already 1/16)
  bool cut() {
    if (disp <= 1.0 / (kDiv * kDiv) || level + kDiv > kRows) return false;
    const int lv = level - 1, top = lv + kDiv;
    const double a = st[lv][0], b = st[lv][1], inc = (b - a) / kDiv;
    double start = a;
    for (int i = top; i >= lv + 1; --i) {
      st[i][0] = start;
      st[i][1] = start + inc;
      st[i][2] = inc;
      start = start + inc;
    }
    st[top][3] = st[lv][3] / kDiv;
    st[lv][3] = st[lv][3] / st[top][3];
    for (int i = lv + 1; i < top; ++i) st[i][3] = 1.0;
    level = top + 1;
    scaling = st[top][3];
    disp = st[top][2];
    return true;
  }
  // process type 2: after a converged sub-increment (true: another
  // follows)
  bool next() {
    if (level <= 1) return false;
    if (st[level - 1][1] == st[level - 2][1]) {
      level -= 2;
      if (level == 0) {
        scaling = st[0][3];
        level = 1;
        return false;
      }
      st[level - 1][3] = kDiv * st[level - 1][3];
      scaling = st[level - 1][3];
    } else {
      level -= 1;
    }
    if (level <= 1) return false;
    scaling = st[level - 1][3];
    disp = st[level - 1][2];
    return true;
  }
};

struct w3d_problem {
  w3d_ctx* c = nullptr;
  w3d_newton_opts o{};
  int neq = 0, ncdof = 0, nldof = 0;
  int64_t nd = 0;
  double *d_u = nullptr, *d_du = nullptr, *d_ifv = nullptr;
  double* d_k = nullptr;  // assembled diagonal, then coefficients
  double *d_rhs = nullptr, *d_x = nullptr, *d_r = nullptr, *d_p = nullptr,
         *d_q = nullptr;
  double* d_cimp = nullptr;  // a full step's imposed increments (cdof order)
  int32_t *d_eqdof = nullptr, *d_cdof = nullptr, *d_ldof = nullptr;
  int64_t bytes = 0;
  int status = W3D_PROBLEM_RUNNING;
  // the next recovery's iteration (0: the start of an increment, as the
  // reference's iteration 0) and whether its increment is extrapolated
  int step = 1, iter = 0;
  bool extrap = true;
  double dt0 = 1.0;  // the context's time increment of a whole step
  // what the next advance does before its recovery
  enum Pending { kNone, kCommitStep, kCommitSub } pending = kNone;
  AdaptStack ad;
  int iters_step = 0, pcg_iters_step = 0, cuts_step = 0;
  double res = 0.0, react = 0.0;
  struct Record {
    int newton = 0, pcg = 0, cuts = 0;
    double load = 0.0, res = 0.0, react = 0.0;
  };
  std::vector<Record> records;
};

struct w3d_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int n_cu = 0;
  int mm10_grid = 0;
  int pcg_chunk = 8;
  // device tables and their pinned host images, grown on demand
  Mm10Slot *d_slots = nullptr, *h_slots = nullptr;
  PcgSys *d_pcg = nullptr, *h_pcg = nullptr;
  ResSys *d_res = nullptr, *h_res = nullptr;
  int cap_sys = 0;
  double* d_part = nullptr;
  int64_t cap_part = 0;
  int32_t *d_state = nullptr, *h_state = nullptr;
  double *d_sc = nullptr, *d_out = nullptr, *h_out = nullptr;
  int* d_counter = nullptr;
  // device time per phase of the advances (w3d_batch_phase_times):
  // 0 recovery (all passes), 1 of which the batched mm10, 2 residual,
  // 3 tangents and assemblies, 4 the batched solve, 5 the update
  static constexpr int kPhases = 6;
  hipEvent_t ev[kPhases][2] = {};
  bool ev_used[kPhases] = {};
  double phase_ms[kPhases] = {};
  int64_t advances = 0, pcg_iterations = 0;
  void mark(int ph, int k) {
    if (ev[ph][k]) {
      (void)hipEventRecord(ev[ph][k], stream);
      if (k == 1) ev_used[ph] = true;
    }
  }
};

namespace {

void free_problem(w3d_problem* p) {
  dfree(p->d_u); dfree(p->d_du); dfree(p->d_ifv); dfree(p->d_k);
  dfree(p->d_rhs); dfree(p->d_x); dfree(p->d_r); dfree(p->d_p); dfree(p->d_q);
  dfree(p->d_cimp);
  dfree(p->d_eqdof); dfree(p->d_cdof); dfree(p->d_ldof);
}

// du for the start of an increment: the free entries scaled by sf (0: no
// extrapolation), the constrained ones the imposed values times disp; the
// time increment scaled by disp as well (mnralg.f:306-309: a sub-increment
// takes its share of the step's time, so rates stay those of the step)
int start_increment(w3d_problem* p, double sf, double disp, hipStream_t s) {
  p->c->dt = p->dt0 * disp;
  int rc = launch_start_increment(p->d_eqdof, p->neq, p->d_cdof, p->d_cimp,
                                  p->ncdof, sf, disp, p->d_du, s);
  return rc ? fail(rc, "w3d_batch_advance: increment launch") : W3D_OK;
}

template <typename T>
void hfree(T*& p) {
  if (p) (void)hipHostFree((void*)p);
  p = nullptr;
}

template <typename T>
int halloc(T** p, size_t n) {
  *p = nullptr;
  if (hipHostMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T),
                    hipHostMallocDefault) != hipSuccess)
    return fail(W3D_ERR_NOMEM, "hipHostMalloc failed");
  return W3D_OK;
}

void free_batch_tables(w3d_batch* b) {
  dfree(b->d_slots); hfree(b->h_slots);
  dfree(b->d_pcg); hfree(b->h_pcg);
  dfree(b->d_res); hfree(b->h_res);
  dfree(b->d_state); hfree(b->h_state);
  dfree(b->d_sc); dfree(b->d_out); hfree(b->h_out);
  b->cap_sys = 0;
}

int reserve(w3d_batch* b, int nsys, int64_t part_words) {
  int rc;
  if (nsys > b->cap_sys) {
    free_batch_tables(b);
    const int cap = std::max(nsys, 2 * b->cap_sys + 8);
    if ((rc = dalloc(&b->d_slots, cap)) || (rc = halloc(&b->h_slots, cap)) ||
        (rc = dalloc(&b->d_pcg, cap)) || (rc = halloc(&b->h_pcg, cap)) ||
        (rc = dalloc(&b->d_res, cap)) || (rc = halloc(&b->h_res, cap)) ||
        (rc = dalloc(&b->d_state, 2 * (size_t)cap)) ||
        (rc = halloc(&b->h_state, 2 * (size_t)cap)) ||
        (rc = dalloc(&b->d_sc, 2 * (size_t)cap)) ||
        (rc = dalloc(&b->d_out, 8 * (size_t)cap)) ||
        (rc = halloc(&b->h_out, 8 * (size_t)cap)))
      return rc;
    b->cap_sys = cap;
  }
  if (part_words > b->cap_part) {
    dfree(b->d_part);
    const int64_t cap = std::max(part_words, 2 * b->cap_part);
    if ((rc = dalloc(&b->d_part, cap))) return rc;
    b->cap_part = cap;
  }
  return W3D_OK;
}

inline int tiles(int64_t n) {
  return (int)std::max<int64_t>(1, (n + kTileRows - 1) / kTileRows);
}

// Routes a context's launches to the batch's stream for one advance.
struct StreamSwap {
  std::vector<std::pair<w3d_ctx*, hipStream_t>> saved;
  void add(w3d_ctx* c, hipStream_t s) {
    saved.push_back({c, c->stream});
    c->stream = s;
  }
  ~StreamSwap() {
    for (auto& cs : saved) cs.first->stream = cs.second;
  }
};

}  // namespace

extern "C" {

int w3d_problem_create(w3d_ctx* c, const int32_t* cstmap, const double* du0,
                       const w3d_newton_opts* o, w3d_problem** out) {
  if (!c || !cstmap || !du0 || !o || !out)
    return fail(W3D_ERR_ARG, "w3d_problem_create: null argument");
  *out = nullptr;
  if (c->asym)
    return fail(W3D_ERR_UNSUPPORTED,
                "w3d_problem_create: asymmetric assembly (Jacobi-PCG needs "
                "the symmetric matrix)");
  if (o->nsteps < 1 || o->max_iters < 1 || o->pcg_maxit < 1 ||
      !(o->tol > 0.0) || !(o->tol_abs >= 0.0) || !(o->pcg_rtol > 0.0) ||
      (o->precision != W3D_FP64 && o->precision != W3D_FP32) ||
      (o->adaptive != 0 && o->adaptive != 1))
    return fail(W3D_ERR_ARG, "w3d_problem_create: bad options");
  HIPCHK(hipSetDevice(c->device));
  int32_t neq = 0;
  int64_t nco = 0;
  int rc = w3d_pattern(c, cstmap, &neq, &nco);
  if (rc) return rc;
  w3d_problem* p = new w3d_problem();
  p->c = c;
  p->o = *o;
  p->dt0 = c->dt;
  p->neq = neq;
  p->nd = 3 * (int64_t)c->nonode;
  std::vector<int32_t> eqdof(std::max(neq, 1)), cdof, ldof;
  std::vector<double> cimp;
  for (int64_t d = 0; d < p->nd; ++d) {
    if (cstmap[d] == 0) {
      eqdof[c->dof_eqn_map[d] - 1] = (int32_t)d;
    } else {
      cdof.push_back((int32_t)d);
      cimp.push_back(du0[d]);
      if (du0[d] != 0.0) ldof.push_back((int32_t)d);
    }
  }
  p->ad.init();
  p->ncdof = (int)cdof.size();
  p->nldof = (int)ldof.size();
  const int64_t b0 = t_dalloc_bytes;
  const std::vector<double> du(du0, du0 + p->nd);
  if ((rc = dalloc(&p->d_u, p->nd)) || (rc = upload(&p->d_du, du)) ||
      (rc = dalloc(&p->d_ifv, p->nd)) ||
      (rc = dalloc(&p->d_k, neq + (size_t)nco)) ||
      (rc = dalloc(&p->d_rhs, neq)) || (rc = dalloc(&p->d_x, neq)) ||
      (rc = dalloc(&p->d_r, neq)) || (rc = dalloc(&p->d_p, neq)) ||
      (rc = dalloc(&p->d_q, neq)) || (rc = upload(&p->d_cimp, cimp)) ||
      (rc = upload(&p->d_eqdof, eqdof)) ||
      (rc = upload(&p->d_cdof, cdof)) || (rc = upload(&p->d_ldof, ldof))) {
    free_problem(p);
    delete p;
    return rc;
  }
  if (hipMemset(p->d_u, 0, p->nd * sizeof(double)) != hipSuccess) {
    free_problem(p);
    delete p;
    return fail(W3D_ERR_DEVICE, "w3d_problem_create: hipMemset");
  }
  p->bytes = c->dev_bytes + (t_dalloc_bytes - b0);
  *out = p;
  return W3D_OK;
}

int w3d_problem_destroy(w3d_problem* p) {
  if (!p) return W3D_OK;
  (void)hipSetDevice(p->c->device);
  (void)hipDeviceSynchronize();
  free_problem(p);
  delete p;
  return W3D_OK;
}

int w3d_problem_status(const w3d_problem* p, int32_t* status, int32_t* step,
                       int32_t* iter) {
  if (!p) return fail(W3D_ERR_ARG, "w3d_problem_status: null problem");
  if (status) *status = p->status;
  if (step) *step = p->step;
  if (iter) *iter = p->iter;
  return W3D_OK;
}

int w3d_problem_step_record(const w3d_problem* p, int32_t step,
                            int32_t* newton_iters, int32_t* pcg_iters,
                            int32_t* cuts, double* load, double* residual,
                            double* reaction) {
  if (!p) return fail(W3D_ERR_ARG, "w3d_problem_step_record: null problem");
  if (step < 1 || step > (int32_t)p->records.size())
    return fail(W3D_ERR_ARG, "w3d_problem_step_record: step not converged");
  const w3d_problem::Record& r = p->records[step - 1];
  if (newton_iters) *newton_iters = r.newton;
  if (pcg_iters) *pcg_iters = r.pcg;
  if (cuts) *cuts = r.cuts;
  if (load) *load = r.load;
  if (residual) *residual = r.res;
  if (reaction) *reaction = r.react;
  return W3D_OK;
}

int w3d_problem_get_u(const w3d_problem* p, double* u) {
  if (!p || !u) return fail(W3D_ERR_ARG, "w3d_problem_get_u: null argument");
  HIPCHK(hipSetDevice(p->c->device));
  HIPCHK(hipMemcpy(u, p->d_u, p->nd * sizeof(double), hipMemcpyDeviceToHost));
  if (p->pending != w3d_problem::kNone) {  // the increment is not in u yet
    std::vector<double> du(p->nd);
    HIPCHK(hipMemcpy(du.data(), p->d_du, p->nd * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < p->nd; ++i) u[i] = u[i] + du[i];
  }
  return W3D_OK;
}

int w3d_problem_device_bytes(const w3d_problem* p, int64_t* bytes) {
  if (!p || !bytes) return fail(W3D_ERR_ARG, "w3d_problem_device_bytes");
  *bytes = p->bytes;
  return W3D_OK;
}

int w3d_batch_create(int32_t device, w3d_batch** out) {
  if (!out) return fail(W3D_ERR_ARG, "w3d_batch_create: null argument");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return fail(W3D_ERR_ARG, "w3d_batch_create: bad device");
  HIPCHK(hipSetDevice(device));
  w3d_batch* b = new w3d_batch();
  b->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) !=
          hipSuccess) {
    delete b;
    return fail(W3D_ERR_DEVICE, "w3d_batch_create: device set-up");
  }
  b->n_cu = prop.multiProcessorCount;
  b->mm10_grid = 4 * std::max(b->n_cu, 1);
  for (int ph = 0; ph < w3d_batch::kPhases; ++ph)
    for (int k = 0; k < 2; ++k)
      if (hipEventCreate(&b->ev[ph][k]) != hipSuccess) b->ev[ph][k] = nullptr;
  int rc = dalloc(&b->d_counter, 1);
  if (rc) {
    (void)hipStreamDestroy(b->stream);
    delete b;
    return rc;
  }
  *out = b;
  return W3D_OK;
}

int w3d_batch_destroy(w3d_batch* b) {
  if (!b) return W3D_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  free_batch_tables(b);
  dfree(b->d_part);
  dfree(b->d_counter);
  for (int ph = 0; ph < w3d_batch::kPhases; ++ph)
    for (int k = 0; k < 2; ++k)
      if (b->ev[ph][k]) (void)hipEventDestroy(b->ev[ph][k]);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return W3D_OK;
}

int w3d_batch_phase_times(const w3d_batch* b, double* ms, int64_t* advances,
                          int64_t* pcg_iterations) {
  if (!b) return fail(W3D_ERR_ARG, "w3d_batch_phase_times: null batch");
  if (ms)
    for (int ph = 0; ph < w3d_batch::kPhases; ++ph) ms[ph] = b->phase_ms[ph];
  if (advances) *advances = b->advances;
  if (pcg_iterations) *pcg_iterations = b->pcg_iterations;
  return W3D_OK;
}

int w3d_batch_advance(w3d_batch* b, w3d_problem* const* ps, int32_t n,
                      int32_t* status) {
  if (!b || (n > 0 && !ps)) return fail(W3D_ERR_ARG, "w3d_batch_advance");
  HIPCHK(hipSetDevice(b->device));
  std::vector<w3d_problem*> run;
  for (int k = 0; k < n; ++k) {
    if (!ps[k] || ps[k]->c->device != b->device)
      return fail(W3D_ERR_ARG, "w3d_batch_advance: problem on another device");
    if (ps[k]->status == W3D_PROBLEM_RUNNING ||
        ps[k]->status == W3D_PROBLEM_STEP_DONE)
      run.push_back(ps[k]);
  }
  for (size_t i = 0; i < run.size(); ++i)
    for (size_t j = 0; j < i; ++j)
      if (run[i]->c == run[j]->c)
        return fail(W3D_ERR_ARG, "w3d_batch_advance: a context appears twice");
  const int nr = (int)run.size();
  hipStream_t s = b->stream;
  StreamSwap swap;
  for (w3d_problem* p : run) swap.add(p->c, s);
  int rc = W3D_OK;
  // 1. commits and the starts of steps and sub-increments
  for (w3d_problem* p : run) {
    if (p->pending != w3d_problem::kNone) {
      if ((rc = enqueue_commit(p->c)) ||
          (rc = launch_axpy1(p->nd, p->d_u, p->d_du, s)))
        return fail(rc ? rc : W3D_ERR_DEVICE, "w3d_batch_advance: commit");
      if (p->pending == w3d_problem::kCommitStep) {
        // a new step extrapolates the last converged increment unscaled
        // (the adaptive stack restarts at 1)
        p->step += 1;
        p->ad.init();
        p->iters_step = p->pcg_iters_step = p->cuts_step = 0;
        rc = start_increment(p, 1.0, 1.0, s);
      } else {
        rc = start_increment(p, p->ad.scaling, p->ad.disp, s);
      }
      if (rc) return rc;
      p->pending = w3d_problem::kNone;
    }
    p->status = W3D_PROBLEM_RUNNING;
    if (p->iter >= 1) p->iters_step += 1;
  }
  // residual tiles (also the update's)
  int ntile_res = 0;
  std::vector<int> res_tile0(nr);
  for (int k = 0; k < nr; ++k) {
    res_tile0[k] = ntile_res;
    ntile_res += tiles(std::max({run[k]->neq, run[k]->ncdof, run[k]->nldof}));
  }
  if ((rc = reserve(b, nr, 3 * (int64_t)ntile_res))) return rc;
  for (int ph = 0; ph < w3d_batch::kPhases; ++ph) b->ev_used[ph] = false;
  // 2. recovery: per-problem passes around one mm10 launch
  b->mark(0, 0);
  std::vector<RecoverLists> lists(nr);
  for (int k = 0; k < nr && !rc; ++k)
    rc = recover_lists(run[k]->c, lists[k]);
  for (int k = 0; k < nr && !rc; ++k)
    rc = enqueue_recover_stage(run[k]->c, lists[k], 0, run[k]->step,
                               run[k]->iter, run[k]->extrap, run[k]->d_u,
                               run[k]->d_du, run[k]->d_ifv, false, 1,
                               run[k]->c->nelblk);
  if (rc) return rc;
  int nslot = 0, nitems = 0;
  for (int k = 0; k < nr; ++k) {
    const RecoverLists& L = lists[k];
    if (L.nwm <= 0) continue;
    Mm10Slot& sl = b->h_slots[nslot++];
    sl.M = run[k]->c->dev();
    sl.work = L.d_wm;
    sl.item0 = nitems;
    sl.nwork = L.nwm;
    sl.step = run[k]->step;
    sl.iter0_off = mm10_iter0_off(run[k]->step, run[k]->iter,
                                  run[k]->extrap ? 1 : 0);
    nitems += L.nwm;
  }
  if (nslot > 0) {
    HIPCHK(hipMemcpyAsync(b->d_slots, b->h_slots, nslot * sizeof(Mm10Slot),
                          hipMemcpyHostToDevice, s));
    b->mark(1, 0);
    rc = launch_mm10_batch(b->d_slots, nslot, nitems, b->d_counter,
                           b->mm10_grid, s);
    b->mark(1, 1);
    if (rc) return fail(rc, "w3d_batch_advance: mm10 launch");
  }
  for (int stage = 2; stage <= 3 && !rc; ++stage)
    for (int k = 0; k < nr && !rc; ++k)
      rc = enqueue_recover_stage(run[k]->c, lists[k], stage, run[k]->step,
                                 run[k]->iter, run[k]->extrap, run[k]->d_u,
                                 run[k]->d_du, run[k]->d_ifv, false, 1,
                                 run[k]->c->nelblk);
  if (rc) return rc;
  b->mark(0, 1);
  // 3. residuals and the decisions
  for (int k = 0; k < nr; ++k) {
    w3d_problem* p = run[k];
    ResSys& R = b->h_res[k];
    R.ifv = p->d_ifv;
    R.eqdof = p->d_eqdof;
    R.cdof = p->d_cdof;
    R.ldof = p->d_ldof;
    R.flags = p->c->d_flags;
    R.rhs = p->d_rhs;
    R.du = p->d_du;
    R.x = p->d_x;
    R.neq = p->neq;
    R.ncdof = p->ncdof;
    R.nldof = p->nldof;
    R.tile0 = res_tile0[k];
    R.ntile = (k + 1 < nr ? res_tile0[k + 1] : ntile_res) - res_tile0[k];
    R.solve = 0;
  }
  if (nr > 0) {
    HIPCHK(hipMemcpyAsync(b->d_res, b->h_res, nr * sizeof(ResSys),
                          hipMemcpyHostToDevice, s));
    b->mark(2, 0);
    if ((rc = launch_residual_batch(b->d_res, nr, ntile_res, b->d_part,
                                    b->d_out, s)))
      return fail(rc, "w3d_batch_advance: residual launch");
    b->mark(2, 1);
    HIPCHK(hipMemcpyAsync(b->h_out, b->d_out, 8 * nr * sizeof(double),
                          hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  std::vector<int> solve;
  for (int k = 0; k < nr; ++k) {
    w3d_problem* p = run[k];
    const double* o = b->h_out + 8 * k;
    p->res = std::sqrt(o[0]);
    p->react = std::sqrt(o[1]);
    // flags: [0] negative plastic multiplier (fatal), [1] a material asks
    // for a cut, [3] det F <= 0
    if (o[3] != 0.0) {
      p->status = W3D_PROBLEM_FAILED;
      continue;
    }
    const bool cut = o[4] != 0.0 || o[6] != 0.0;
    // iteration 0 is never accepted: it only forms the increment's first
    // residual and stiffness (mnralg.f's iteration 0)
    if (!cut && p->iter >= 1 &&
        p->res <= p->o.tol * std::max(p->react, p->o.tol_abs)) {
      p->iter = 0;
      p->extrap = true;
      if (p->o.adaptive && p->ad.next()) {
        p->pending = w3d_problem::kCommitSub;  // the next sub-increment
        continue;
      }
      w3d_problem::Record r;
      r.newton = p->iters_step;
      r.pcg = p->pcg_iters_step;
      r.cuts = p->cuts_step;
      r.load = o[2];
      r.res = p->res;
      r.react = p->react;
      p->records.push_back(r);
      if (p->step >= p->o.nsteps) {
        if ((rc = enqueue_commit(p->c)) ||
            (rc = launch_axpy1(p->nd, p->d_u, p->d_du, s)))
          return fail(rc ? rc : W3D_ERR_DEVICE, "w3d_batch_advance: commit");
        p->status = W3D_PROBLEM_DONE;
      } else {
        p->pending = w3d_problem::kCommitStep;
        p->status = W3D_PROBLEM_STEP_DONE;
      }
      continue;
    }
    if (cut || p->iter >= p->o.max_iters) {
      // restore n and restart on the first sub-increment of a split,
      // without extrapolation
      if (p->o.adaptive && p->ad.cut()) {
        if ((rc = enqueue_rollback(p->c))) return rc;
        if ((rc = start_increment(p, 0.0, p->ad.disp, s))) return rc;
        p->cuts_step += 1;
        p->iter = 0;
        p->extrap = false;
      } else {
        p->status = W3D_PROBLEM_FAILED;
      }
      continue;
    }
    p->iter += 1;  // the next recovery follows this iteration's solve
    solve.push_back(k);
  }
  // 4. tangents, assemblies and the batched solve
  const int ns = (int)solve.size();
  if (ns > 0) {
    int ntile = 0, maxit = 0;
    b->mark(3, 0);
    for (int i = 0; i < ns; ++i) {
      w3d_problem* p = run[solve[i]];
      w3d_ctx* c = p->c;
      rc = launch_tangent_mfma(c->dev(), c->d_work_tan, c->tan_groups,
                               c->tan_start, c->tan_count, p->o.precision, s);
      if (!rc)
        rc = c->enqueue_assemble(p->o.precision == W3D_FP32, p->d_k,
                                 p->d_k + p->neq, s);
      if (!rc) rc = enqueue_csr_values(c, p->d_k, p->d_k + p->neq);
      if (rc) return fail(rc, "w3d_batch_advance: tangent/assembly launch");
      c->tangent_precision = p->o.precision;
      PcgSys& S = b->h_pcg[i];
      S.rp = c->d_csr_rp;
      S.col = c->d_csr_col;
      S.val = c->d_csr_val;
      S.diag = p->d_k;
      S.b = p->d_rhs;
      S.x = p->d_x;
      S.r = p->d_r;
      S.p = p->d_p;
      S.q = p->d_q;
      S.rtol = p->o.pcg_rtol;
      S.maxit = p->o.pcg_maxit;
      S.n = p->neq;
      S.tile0 = ntile;
      S.ntile = tiles(p->neq);
      ntile += S.ntile;
      maxit = std::max(maxit, S.maxit);
    }
    b->mark(3, 1);
    // the solve's partial sums follow the residual's in d_part
    if ((rc = reserve(b, nr, 3 * (int64_t)ntile_res + 7 * (int64_t)ntile)))
      return rc;
    double* part = b->d_part + 3 * (int64_t)ntile_res;
    HIPCHK(hipMemcpyAsync(b->d_pcg, b->h_pcg, ns * sizeof(PcgSys),
                          hipMemcpyHostToDevice, s));
    b->mark(4, 0);
    if ((rc = launch_pcg_batch_init(b->d_pcg, ns, ntile, part, b->d_state,
                                    b->d_sc, s)))
      return fail(rc, "w3d_batch_advance: solver launch");
    // every system ends by its own test or its own iteration limit; the
    // host looks at the states once per chunk of iterations
    int it = 0;
    bool all_done = false;
    while (!all_done && it < maxit) {
      const int end = std::min(maxit, it + b->pcg_chunk);
      for (; it < end; ++it)
        if ((rc = launch_pcg_batch_iter(b->d_pcg, ns, ntile, it, part,
                                        b->d_state, b->d_sc, s)))
          return fail(rc, "w3d_batch_advance: solver launch");
      HIPCHK(hipMemcpyAsync(b->h_state, b->d_state, 2 * ns * sizeof(int32_t),
                            hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      all_done = true;
      for (int i = 0; i < ns; ++i) all_done = all_done && b->h_state[2 * i];
    }
    b->mark(4, 1);
    b->pcg_iterations += it;
    for (int i = 0; i < ns; ++i) {
      run[solve[i]]->pcg_iters_step += b->h_state[2 * i + 1];
      b->h_res[solve[i]].solve = 1;
    }
    // 5. du += x
    HIPCHK(hipMemcpyAsync(b->d_res, b->h_res, nr * sizeof(ResSys),
                          hipMemcpyHostToDevice, s));
    b->mark(5, 0);
    if ((rc = launch_du_batch(b->d_res, nr, ntile_res, s)))
      return fail(rc, "w3d_batch_advance: update launch");
    b->mark(5, 1);
  }
  HIPCHK(hipStreamSynchronize(s));
  b->advances += 1;
  for (int ph = 0; ph < w3d_batch::kPhases; ++ph) {
    float t = 0.0f;
    if (b->ev_used[ph] &&
        hipEventElapsedTime(&t, b->ev[ph][0], b->ev[ph][1]) == hipSuccess)
      b->phase_ms[ph] += t;
  }
  if (status)
    for (int k = 0; k < n; ++k) status[k] = ps[k]->status;
  return W3D_OK;
}

}  // extern "C"
