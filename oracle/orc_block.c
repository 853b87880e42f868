/*
 * orc_block.c -- oracle: the block model, the recovery / internal-force
 * driver, the block tangent, automatic blocking and the VSS sparse pattern
 * and row-by-row assembly.
 *
 * TEST INFRASTRUCTURE ONLY (see w3d_oracle.h).  Restates, in the reference's
 * own block loop order and array layouts:
 *   inelbk.f:282-548 (inelbk_simple_blocking, inelbk_chk_match)
 *   setup.f:160-190, 268-276 (inverse incidences, cp/dcp/icp)
 *   reopen.f:1315 (edest_init)
 *   drive_eps_sig_internal_forces.f:20, :399, :1653 (driver, do_nleps_block,
 *     dupstr_blocked), rknstr.f:563 (rknstr_sm_displ), rstgp1.f:351 (rstgp2),
 *     gtlsn1.f:25, rstgp1.f:614 (drive_01_update), :998 (drive_01_update_b),
 *     :5189 (rstgp1_b), :5360 (rstgp1_store_cep), rplstr.f:15,
 *     rknifv.f:141 (rknifv_std), gpifv1.f:16, addifv.f:16
 *   tanstf.f:119, rktstf.f:15, gptns1.f:16, :1174 (extract_D_symmetric),
 *     :1628 (bdbtgen), gpmas1.f:399 (adjust_cnst), rktstf.f:221 (transpose)
 *   assemble_code.f:20, :151, :260, :355, :561, :659, :785
 *
 * The scatter of element forces into the structure vector (addifv) uses
 * OpenMP atomics in the reference, whose summation order is not fixed; the
 * oracle sums each structure dof over its inverse-incidence list in element
 * order, which is one of the orders the reference can take and the order the
 * device library uses.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "w3d_oracle.h"

#define MXND 20
#define MXDOF 60
#define MXGP 27

static __thread char g_err[512] = "no error";
const char* orc_last_error(void) { return g_err; }
static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

typedef struct {
  int span, felem, etype, order, ngp, nnode, totdof, utsz, mat, model,
      hist, bbar;
  /* device-free host state in the reference layouts */
  double *urcs_n, *urcs_n1;   /* (9, ngp, span) */
  double *eps_n, *eps_n1;     /* (6, ngp, span) */
  double *hist_n, *hist_n1;   /* (hist, ngp, span) */
  double *cep;                /* (21, span, ngp) */
  double *ek;                 /* (utsz, span) */
  double *einfvec;            /* (span, totdof) */
  double *evol;               /* (span) */
  int32_t* edest;             /* (totdof, span) 1-based structure dofs */
  int ncry;                   /* mm10: crystals per point                */
  orc_crystal* crys;          /* mm10: (ncry, span) crystal constants    */
  int geonl;                  /* finite strains and rotations            */
  double *rot_n, *rot_n1;     /* geonl: (9, ngp, span) [R] column-major  */
  double* ekf;                /* asymmetric: (T*T, span) full Ke         */
  double energy, pwork;       /* the last recovery's block sums (rstgp1_b) */
} orc_block;

struct orc_model {
  int nonode, noelem, nelblk, nmat;
  double* c;
  int32_t *incid, *incmap, *etype, *order, *matnum, *bbar, *geonl;
  /* displacements of the last recovery (the tangent's configuration) */
  double *u_last, *du_last;
  w3d_material* mats;
  double eps_bbar;
  int32_t* elblks; /* (4, nelblk) */
  /* skewed constraints: trn (nonode) and trnmat (9 per node, column-major
     mat(r, c) at r + 3c), NULL when no node has one */
  int32_t* trn;
  double* trnmat;
  double dt;       /* time increment of the load step */
  /* temperatures of the load step (main_data temper_nodes, dtemp_nodes,
     temper_elems, dtemp_elems; NULL = zero) and the `temperatures` flag */
  double *tn, *dtn, *te, *dte;
  int temps_incr;
  int asym;        /* asymmetric assembly                */
  orc_block* blk;
  int32_t* elem_blk; /* (noelem) 0-based block of each element */
  /* inverse incidences: per node, elements in ascending order */
  int32_t *inv_ptr, *inv_list;
  /* sparsity */
  int neqns;
  int64_t ncoeff;
  int32_t *dof_eqn_map, *eqn_node_map, *k_ptrs, *k_indexes;
  int64_t* row_start;
};

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }

void orc_set_time_increment(orc_model* om, double dt) { om->dt = dt; }

static double* copy_or_null(double* old, const double* src, int64_t n) {
  free(old);
  if (!src) return NULL;
  double* d = (double*)malloc(sizeof(double) * n);
  memcpy(d, src, sizeof(double) * n);
  return d;
}

int orc_set_temperatures(orc_model* om, const double* temper_nodes,
                         const double* dtemp_nodes, const double* temper_elems,
                         const double* dtemp_elems, int incremental) {
  om->tn = copy_or_null(om->tn, temper_nodes, om->nonode);
  om->dtn = copy_or_null(om->dtn, dtemp_nodes, om->nonode);
  om->te = copy_or_null(om->te, temper_elems, om->noelem);
  om->dte = copy_or_null(om->dte, dtemp_elems, om->noelem);
  om->temps_incr = incremental != 0;
  return 0;
}

static int ncry_of(const w3d_material* m) {
  return m->ncrystals > 0 ? m->ncrystals : 1;
}
static int hist_size(const w3d_material* m) {
  return m->model == W3D_MM01 ? 11 : orc_mm10_hist_size(ncry_of(m));
}

/* ------------------------------------------------------------------ */
/* blocking: inelbk_simple_blocking / inelbk_chk_match                 */
/* ------------------------------------------------------------------ */
static int compatible(const orc_model* om, int a, int b) {
  if (om->etype[a] != om->etype[b]) return 0;
  if (om->matnum[a] != om->matnum[b]) return 0;
  if (om->order[a] != om->order[b]) return 0;
  if (orc_ngp(om->etype[a], om->order[a]) !=
      orc_ngp(om->etype[b], om->order[b]))
    return 0;
  if (om->bbar[a] != om->bbar[b]) return 0;
  if (om->geonl[a] != om->geonl[b]) return 0;
  return 1;
}

static int auto_blocking(orc_model* om, int auto_size) {
  int cap = om->noelem + 1;
  int32_t* t = (int32_t*)malloc(sizeof(int32_t) * 4 * cap);
  int nb = 1, cur = 1, felem = 1;
  t[0] = 1; /* span */
  t[1] = 1; /* felem */
  t[2] = 0;
  t[3] = 0;
  for (int e = 2; e <= om->noelem; ++e) {
    int newblk = 0;
    cur = cur + 1;
    if (cur > auto_size) newblk = 1;
    if (!newblk && !compatible(om, felem - 1, e - 1)) newblk = 1;
    if (!newblk) {
      t[4 * (nb - 1) + 0] = cur;
      continue;
    }
    nb += 1;
    felem = e;
    t[4 * (nb - 1) + 0] = 1;
    t[4 * (nb - 1) + 1] = felem;
    t[4 * (nb - 1) + 2] = 0;
    t[4 * (nb - 1) + 3] = 0;
    cur = 1;
  }
  om->nelblk = nb;
  om->elblks = t;
  return 0;
}

/* ------------------------------------------------------------------ */
/* model creation                                                     */
/* ------------------------------------------------------------------ */
orc_model* orc_create(const w3d_model* m) {
  if (!m || m->abi_version != W3D_ABI_VERSION) {
    fail(W3D_ERR_ARG, "orc_create: bad model or ABI version");
    return NULL;
  }
  orc_model* om = (orc_model*)calloc(1, sizeof(orc_model));
  om->nonode = m->nonode;
  om->noelem = m->noelem;
  om->nmat = m->nmat;
  om->eps_bbar = m->eps_bbar;
  om->dt = 1.0;
  om->c = (double*)malloc(sizeof(double) * 3 * m->nonode);
  memcpy(om->c, m->c, sizeof(double) * 3 * m->nonode);
  om->etype = (int32_t*)malloc(sizeof(int32_t) * m->noelem);
  om->order = (int32_t*)malloc(sizeof(int32_t) * m->noelem);
  om->matnum = (int32_t*)malloc(sizeof(int32_t) * m->noelem);
  om->bbar = (int32_t*)malloc(sizeof(int32_t) * m->noelem);
  om->geonl = (int32_t*)malloc(sizeof(int32_t) * m->noelem);
  om->u_last = (double*)calloc(3 * (size_t)m->nonode, sizeof(double));
  om->du_last = (double*)calloc(3 * (size_t)m->nonode, sizeof(double));
  if (m->trn && m->trnmat) {
    om->trn = (int32_t*)malloc(sizeof(int32_t) * m->nonode);
    om->trnmat = (double*)malloc(sizeof(double) * 9 * m->nonode);
    memcpy(om->trn, m->trn, sizeof(int32_t) * m->nonode);
    memcpy(om->trnmat, m->trnmat, sizeof(double) * 9 * m->nonode);
  }
  om->incmap = (int32_t*)malloc(sizeof(int32_t) * m->noelem);
  int64_t ninc = 0;
  for (int e = 0; e < m->noelem; ++e) {
    int nn = orc_nnode(m->etype[e]);
    if (nn == 0 || orc_ngp(m->etype[e], m->int_order[e]) == 0) {
      fail(W3D_ERR_UNSUPPORTED, "orc_create: element type/order");
      orc_destroy(om);
      return NULL;
    }
    ninc += nn;
  }
  om->incid = (int32_t*)malloc(sizeof(int32_t) * ninc);
  int64_t pos = 0;
  for (int e = 0; e < m->noelem; ++e) {
    int nn = orc_nnode(m->etype[e]);
    om->etype[e] = m->etype[e];
    om->order[e] = m->int_order[e];
    om->matnum[e] = m->matnum[e];
    om->bbar[e] = (m->etype[e] == W3D_L3DISOP) && m->bbar && m->bbar[e];
    om->geonl[e] = m->geonl && m->geonl[e] ? 1 : 0;
    om->incmap[e] = (int32_t)pos + 1;
    for (int j = 0; j < nn; ++j)
      om->incid[pos + j] = m->incid[m->incmap[e] - 1 + j];
    pos += nn;
  }
  om->mats = (w3d_material*)malloc(sizeof(w3d_material) * m->nmat);
  memcpy(om->mats, m->mats, sizeof(w3d_material) * m->nmat);
  for (int e = 0; e < m->noelem; ++e) {
    if (om->matnum[e] < 1 || om->matnum[e] > m->nmat) {
      fail(W3D_ERR_ARG, "orc_create: bad material number");
      orc_destroy(om);
      return NULL;
    }
    const w3d_material* mt = &om->mats[om->matnum[e] - 1];
    int ok = mt->model == W3D_MM01;
    if (mt->model == W3D_MM10)
      ok = (mt->slip_type == 1 || mt->slip_type == 2) &&
           (mt->elastic_type == 1 || mt->elastic_type == 2) &&
           mt->angle_convention == 1 &&
           (mt->angle_type == 1 || mt->angle_type == 2) &&
           ncry_of(mt) <= (m->angle_stride > 0 ? m->angle_stride : 1) &&
           m->angles != NULL;
    if (!ok) {
      fail(W3D_ERR_UNSUPPORTED, "orc_create: material model");
      orc_destroy(om);
      return NULL;
    }
  }
  /* blocking */
  if (m->nelblk > 0) {
    om->nelblk = m->nelblk;
    om->elblks = (int32_t*)malloc(sizeof(int32_t) * 4 * m->nelblk);
    memcpy(om->elblks, m->elblks, sizeof(int32_t) * 4 * m->nelblk);
    int next = 1;
    for (int b = 0; b < m->nelblk; ++b) {
      int span = om->elblks[4 * b], felem = om->elblks[4 * b + 1];
      if (span < 1 || felem != next) {
        fail(W3D_ERR_ARG, "orc_create: blocks must tile the elements");
        orc_destroy(om);
        return NULL;
      }
      for (int i = 1; i < span; ++i)
        if (!compatible(om, felem - 1, felem - 1 + i)) {
          fail(W3D_ERR_ARG, "orc_create: incompatible elements in block");
          orc_destroy(om);
          return NULL;
        }
      next = felem + span;
    }
    if (next != m->noelem + 1) {
      fail(W3D_ERR_ARG, "orc_create: blocks must tile the elements");
      orc_destroy(om);
      return NULL;
    }
  } else {
    int as = m->auto_size > 0 ? m->auto_size : 128;
    auto_blocking(om, as);
  }
  /* blocks */
  om->blk = (orc_block*)calloc(om->nelblk, sizeof(orc_block));
  om->elem_blk = (int32_t*)malloc(sizeof(int32_t) * m->noelem);
  for (int b = 0; b < om->nelblk; ++b) {
    orc_block* B = &om->blk[b];
    B->span = om->elblks[4 * b];
    B->felem = om->elblks[4 * b + 1];
    int fe = B->felem - 1;
    B->etype = om->etype[fe];
    B->order = om->order[fe];
    B->ngp = orc_ngp(B->etype, B->order);
    B->nnode = orc_nnode(B->etype);
    B->totdof = 3 * B->nnode;
    B->utsz = B->totdof * (B->totdof + 1) / 2;
    B->mat = om->matnum[fe] - 1;
    B->model = om->mats[B->mat].model;
    B->hist = hist_size(&om->mats[B->mat]);
    B->bbar = om->bbar[fe];
    B->geonl = om->geonl[fe];
    int64_t nq = (int64_t)B->ngp * B->span;
    if (B->geonl) { /* rotation_init (reopen.f:1840): unit matrices */
      B->rot_n = (double*)calloc(9 * nq, sizeof(double));
      B->rot_n1 = (double*)calloc(9 * nq, sizeof(double));
      for (int64_t q = 0; q < nq; ++q) {
        B->rot_n[9 * q] = B->rot_n[9 * q + 4] = B->rot_n[9 * q + 8] = 1.0;
        B->rot_n1[9 * q] = B->rot_n1[9 * q + 4] = B->rot_n1[9 * q + 8] = 1.0;
      }
    }
    B->urcs_n = (double*)calloc(9 * nq, sizeof(double));
    B->urcs_n1 = (double*)calloc(9 * nq, sizeof(double));
    B->eps_n = (double*)calloc(6 * nq, sizeof(double));
    B->eps_n1 = (double*)calloc(6 * nq, sizeof(double));
    B->hist_n = (double*)calloc(B->hist * nq, sizeof(double));
    B->hist_n1 = (double*)calloc(B->hist * nq, sizeof(double));
    B->cep = (double*)calloc(21 * nq, sizeof(double));
    B->ek = (double*)calloc((int64_t)B->utsz * B->span, sizeof(double));
    B->einfvec = (double*)calloc((int64_t)B->totdof * B->span, sizeof(double));
    B->evol = (double*)calloc(B->span, sizeof(double));
    B->edest = (int32_t*)malloc(sizeof(int32_t) * B->totdof * B->span);
    if (B->model == W3D_MM10) {
      const w3d_material* mt = &om->mats[B->mat];
      int stride = m->angle_stride > 0 ? m->angle_stride : 1;
      B->ncry = ncry_of(mt);
      B->crys = (orc_crystal*)malloc(sizeof(orc_crystal) * B->ncry * B->span);
      for (int i = 0; i < B->span; ++i)
        for (int c = 0; c < B->ncry; ++c)
          orc_mm10_crystal(mt, m->angles + 3 * ((int64_t)stride * (fe + i) + c),
                           &B->crys[i * B->ncry + c]);
    }
    for (int i = 0; i < B->span; ++i) {
      int e = fe + i;
      om->elem_blk[e] = b;
      const int32_t* nodes = om->incid + om->incmap[e] - 1;
      for (int j = 0; j < B->nnode; ++j) {
        int d = 3 * (nodes[j] - 1) + 1;
        B->edest[i * B->totdof + j] = d;
        B->edest[i * B->totdof + B->nnode + j] = d + 1;
        B->edest[i * B->totdof + 2 * B->nnode + j] = d + 2;
      }
    }
  }
  /* inverse incidences (setup.f:160-190): repeated nodes of an element
     are listed once */
  om->inv_ptr = (int32_t*)calloc(m->nonode + 1, sizeof(int32_t));
  for (int e = 0; e < m->noelem; ++e) {
    int nn = orc_nnode(om->etype[e]);
    const int32_t* nodes = om->incid + om->incmap[e] - 1;
    for (int j = 0; j < nn; ++j) {
      int rep = 0;
      for (int k = 0; k < j; ++k) rep |= nodes[k] == nodes[j];
      if (!rep) om->inv_ptr[nodes[j]] += 1;
    }
  }
  for (int n = 0; n < m->nonode; ++n) om->inv_ptr[n + 1] += om->inv_ptr[n];
  om->inv_list = (int32_t*)malloc(sizeof(int32_t) * om->inv_ptr[m->nonode]);
  int32_t* fillp = (int32_t*)malloc(sizeof(int32_t) * m->nonode);
  for (int n = 0; n < m->nonode; ++n) fillp[n] = om->inv_ptr[n];
  for (int e = 0; e < m->noelem; ++e) {
    int nn = orc_nnode(om->etype[e]);
    const int32_t* nodes = om->incid + om->incmap[e] - 1;
    for (int j = 0; j < nn; ++j) {
      int rep = 0;
      for (int k = 0; k < j; ++k) rep |= nodes[k] == nodes[j];
      if (!rep) om->inv_list[fillp[nodes[j] - 1]++] = e + 1;
    }
  }
  free(fillp);
  return om;
}

void orc_destroy(orc_model* om) {
  if (!om) return;
  if (om->blk)
    for (int b = 0; b < om->nelblk; ++b) {
      orc_block* B = &om->blk[b];
      free(B->urcs_n); free(B->urcs_n1); free(B->eps_n); free(B->eps_n1);
      free(B->hist_n); free(B->hist_n1); free(B->cep); free(B->ek);
      free(B->einfvec); free(B->evol); free(B->edest); free(B->crys);
      free(B->ekf);
      free(B->rot_n); free(B->rot_n1);
    }
  free(om->blk); free(om->elem_blk); free(om->c); free(om->incid);
  free(om->incmap); free(om->etype); free(om->order); free(om->matnum);
  free(om->bbar); free(om->geonl); free(om->u_last); free(om->du_last);
  free(om->trn); free(om->trnmat);
  free(om->tn); free(om->dtn); free(om->te); free(om->dte);
  free(om->mats); free(om->elblks); free(om->inv_ptr);
  free(om->inv_list); free(om->dof_eqn_map); free(om->eqn_node_map);
  free(om->k_ptrs); free(om->k_indexes); free(om->row_start);
  free(om);
}

int orc_get_blocking(const orc_model* om, int32_t* nelblk, int32_t* elblks) {
  if (nelblk) *nelblk = om->nelblk;
  if (elblks) memcpy(elblks, om->elblks, sizeof(int32_t) * 4 * om->nelblk);
  return 0;
}

int orc_block_info(const orc_model* om, int blk, int32_t* info5) {
  if (blk < 1 || blk > om->nelblk) return fail(W3D_ERR_ARG, "bad block");
  const orc_block* B = &om->blk[blk - 1];
  info5[0] = B->span;
  info5[1] = B->ngp;
  info5[2] = B->hist;
  info5[3] = B->totdof;
  info5[4] = B->utsz;
  return 0;
}

/* ------------------------------------------------------------------ */
/* shared per-block kinematics                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  double w[MXGP];
  double nxi[MXGP][MXND], neta[MXGP][MXND], nzeta[MXGP][MXND];
  double sf[MXGP][MXND]; /* shape functions (gauss_pt_temps) */
} gp_table;

static void make_table(const orc_block* B, gp_table* t) {
  for (int g = 0; g < B->ngp; ++g) {
    double xi, eta, zeta;
    orc_getgpts(B->etype, B->order, g + 1, &xi, &eta, &zeta, &t->w[g]);
    orc_derivs(B->etype, xi, eta, zeta, t->nxi[g], t->neta[g], t->nzeta[g]);
    orc_shapef(B->etype, xi, eta, zeta, t->sf[g]);
  }
}

/* Temperatures of element i's nodes (dupstr_blocked,
   drive_eps_sig_internal_forces.f:1756-1790): gadtemps (gatemps.f:63)
   node + element increments when the step imposes temperature changes,
   else zero; gatemps (:117) totals at the end of the step. */
static void elem_temps(const orc_model* om, const orc_block* B, int i,
                       double* dtn, double* tn1) {
  const int elem = B->felem + i; /* 1-based */
  const int32_t* inc = om->incid + om->incmap[elem - 1] - 1;
  for (int j = 0; j < B->nnode; ++j) {
    const int node = inc[j];
    dtn[j] = om->temps_incr ? (om->dtn ? om->dtn[node - 1] : 0.0) +
                                  (om->dte ? om->dte[elem - 1] : 0.0)
                            : 0.0;
    tn1[j] = (om->tn ? om->tn[node - 1] : 0.0) +
             (om->te ? om->te[elem - 1] : 0.0) + dtn[j];
  }
}

/* gauss_pt_temps (gp_temp_eps.f:17): sum of shape function x nodal value,
   from zero, in node order */
static double gp_interp(int nnode, const double* sf, const double* v) {
  double s = 0.0;
  for (int j = 0; j < nnode; ++j) s = s + sf[j] * v[j];
  return s;
}

/* linear_interpolate (setup_mat_props.f:361) */
static double lin_interp(double x, int n, const double* xs,
                         const double* ys, int ystride) {
  if (x <= xs[0]) return ys[0];
  if (x >= xs[n - 1]) return ys[(n - 1) * ystride];
  for (int p = 1; p < n; ++p) {
    if (x > xs[p]) continue;
    double x1 = xs[p - 1], x2 = xs[p], y1 = ys[(p - 1) * ystride],
           y2 = ys[p * ystride];
    return y1 + (x - x1) * (y2 - y1) / (x2 - x1);
  }
  return ys[(n - 1) * ystride];
}

/* mm01 properties of one point: setup_mm01_rknstr (rknstr.f:1018; REAL
   props, H' in REAL), replaced by the segmental curve set where the
   material has one: set_up_segmental (setup_mat_props.f:67; e, nu and the
   isotropic alpha at the point's temperatures at n+1 and n, the curve's
   stresses at n+1) and set_up_h_prime (:406). */
typedef struct {
  double e, nu, e_n, nu_n, hp, yld, alpha[6], alpha_n[6];
} pt_props;

static void point_props(const w3d_material* mat, double t_n1, double dtemp,
                        pt_props* p) {
  p->e = p->e_n = (double)mat->e;
  p->nu = p->nu_n = (double)mat->nu;
  p->hp = (double)orc_hprime(mat->e, mat->tan_e);
  p->yld = (double)mat->yld_pt;
  for (int k = 0; k < 6; ++k) p->alpha[k] = p->alpha_n[k] = mat->alpha[k];
  if (mat->curve_type == 0) return;
  double s1 = mat->curve_sig[0][0], s2 = mat->curve_sig[0][1];
  if (mat->curve_type == 2) {
    const int n = mat->ncurves;
    const double* ts = mat->curve_temp;
    const double t_n = t_n1 - dtemp;
    p->e = lin_interp(t_n1, n, ts, mat->curve_e, 1);
    p->e_n = lin_interp(t_n, n, ts, mat->curve_e, 1);
    p->nu = lin_interp(t_n1, n, ts, mat->curve_nu, 1);
    p->nu_n = lin_interp(t_n, n, ts, mat->curve_nu, 1);
    double a = lin_interp(t_n1, n, ts, mat->curve_alpha, 1);
    p->alpha[0] = p->alpha[1] = p->alpha[2] = a;
    a = lin_interp(t_n, n, ts, mat->curve_alpha, 1);
    p->alpha_n[0] = p->alpha_n[1] = p->alpha_n[2] = a;
    s1 = lin_interp(t_n1, n, ts, &mat->curve_sig[0][0], 2);
    s2 = lin_interp(t_n1, n, ts, &mat->curve_sig[0][1], 2);
  }
  double dsigma = s2 - s1;
  double depspls = mat->curve_eps[1] - mat->curve_eps[0];
  p->hp = dsigma / depspls;
  p->yld = s1;
}

/* gp_temp_eps (gp_temp_eps.f:163): the (negative) thermal strain
   increment added to the displacement increment deps */
static void thermal_strain(const double* alpha_n1, const double* alpha_n,
                           double alpha_zero, double t_n1, double dtemp,
                           double* deps) {
  double temps_n = t_n1 - dtemp;
  double tn = temps_n - alpha_zero, tn1 = t_n1 - alpha_zero;
  for (int k = 0; k < 6; ++k) {
    double th = alpha_n1[k] * tn1 - alpha_n[k] * tn;
    double ut = 0.0 - th;
    deps[k] = deps[k] + ut;
  }
}

/* ce (3N: x-block, y-block, z-block) of element i of block B */
static void gather_coords(const orc_model* om, const orc_block* B, int i,
                          double* ce) {
  const int32_t* ed = B->edest + (int64_t)i * B->totdof;
  for (int k = 0; k < B->totdof; ++k) ce[k] = om->c[ed[k] - 1];
}

/* Skewed constraints of element i's nodes: mats[j] = trnmat of local node
   j (column-major 3x3) or NULL (duptrans, tanstf.f:733); returns whether
   any node has one (trn_e_flags). */
static int elem_trn(const orc_model* om, const orc_block* B, int i,
                    const double** mats) {
  int any = 0;
  const int32_t* ed = B->edest + (int64_t)i * B->totdof;
  for (int j = 0; j < B->nnode; ++j) {
    mats[j] = NULL;
    if (!om->trn) continue;
    const int node = (ed[j] - 1) / 3; /* edest of component x of node j */
    if (om->trn[node]) {
      mats[j] = om->trnmat + 9 * (int64_t)node;
      any = 1;
    }
  }
  return any;
}

/* trnvec / trnvecs (trnvec.f:17,109) on an element vector in the local
   dof order (component c of node j at c N + j): switch 1 v = T v (global
   to constraint compatible), switch 2 v = T^T v. */
static void trnvec(double* v, int N, const double** mats, int sw) {
  for (int j = 0; j < N; ++j) {
    const double* t = mats[j];
    if (!t) continue;
    double x = v[j], y = v[N + j], z = v[2 * N + j], r[3];
    for (int a = 0; a < 3; ++a) {
      if (sw == 1)
        r[a] = t[a + 0] * x + t[a + 3] * y + t[a + 6] * z;
      else
        r[a] = t[0 + 3 * a] * x + t[1 + 3 * a] * y + t[2 + 3 * a] * z;
    }
    v[j] = r[0];
    v[N + j] = r[1];
    v[2 * N + j] = r[2];
  }
}

/* trnmtx (trnmtx.f:19): Ke = T Ke T^T on the full T x T matrix a (column
   major, a[r + T c]), rows first, then columns, node by node. */
static void trnmtx(double* a, int N, const double** mats) {
  const int T = 3 * N;
  for (int j = 0; j < N; ++j) { /* rows j, N+j, 2N+j <- mat . rows */
    const double* t = mats[j];
    if (!t) continue;
    for (int col = 0; col < T; ++col) {
      double r0 = a[j + T * col], r1 = a[N + j + T * col],
             r2 = a[2 * N + j + T * col], s3[3];
      for (int r = 0; r < 3; ++r)
        s3[r] = t[r + 0] * r0 + t[r + 3] * r1 + t[r + 6] * r2;
      a[j + T * col] = s3[0];
      a[N + j + T * col] = s3[1];
      a[2 * N + j + T * col] = s3[2];
    }
  }
  for (int j = 0; j < N; ++j) { /* columns j, N+j, 2N+j <- cols . mat^T */
    const double* t = mats[j];
    if (!t) continue;
    for (int row = 0; row < T; ++row) {
      double c0 = a[row + T * j], c1 = a[row + T * (N + j)],
             c2 = a[row + T * (2 * N + j)], s3[3];
      for (int r = 0; r < 3; ++r)
        s3[r] = t[r + 0] * c0 + t[r + 3] * c1 + t[r + 6] * c2;
      a[row + T * j] = s3[0];
      a[row + T * (N + j)] = s3[1];
      a[row + T * (2 * N + j)] = s3[2];
    }
  }
}

/* Jacobians at every point of one element, then (hex8 B-bar) the
   volume-averaged derivatives: rknstr_sm_displ / rktstf / rknifv_std. */
static void elem_kinematics(const orc_block* B, const gp_table* t,
                            const double* ce, double dj[MXGP],
                            double gama[MXGP][9], double vol[24]) {
  double volume = 0.0;
  if (B->bbar) memset(vol, 0, sizeof(double) * 24);
  for (int g = 0; g < B->ngp; ++g) {
    double jac[9];
    orc_jacob1(B->nnode, t->nxi[g], t->neta[g], t->nzeta[g], ce, jac, &dj[g],
               gama[g]);
    if (B->bbar) { /* vol_terms, vol_terms.f:14 */
      const double* a = gama[g];
      double d = dj[g];
      for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 8; ++j)
          vol[3 * j + k] = vol[3 * j + k] + (a[3 * k] * t->nxi[g][j] +
                                             a[3 * k + 1] * t->neta[g][j] +
                                             a[3 * k + 2] * t->nzeta[g][j]) *
                                                d;
      volume = volume + d;
    }
  }
  if (B->bbar) { /* vol_avg, vol_avg.f:13 */
    double voli = 1.0 / volume;
    for (int k = 0; k < 3; ++k)
      for (int j = 0; j < 8; ++j) vol[3 * j + k] = vol[3 * j + k] * voli;
  }
}

/* ------------------------------------------------------------------ */
/* recovery: do_nleps_block for one block                              */
/* ------------------------------------------------------------------ */
static int recover_block(orc_model* om, int b, int step, int iter,
                         int extrapolated, const double* u, const double* du,
                         double* energy, double* pwork) {
  (void)u; /* small-displacement form: u enters only through du */
  orc_block* B = &om->blk[b];
  const w3d_material* mat = &om->mats[B->mat];
  gp_table t;
  make_table(B, &t);
  const int S = B->span, G = B->ngp, N = B->nnode, T = B->totdof,
            H = B->hist;
  const double scalar = B->etype == W3D_TET10 ? 1.0 / 6.0 : 1.0;
  /* material property vectors: setup_mm01_rknstr (rknstr.f:1018), per
     point where temperature-dependent curves interpolate them */
  const double beta = (double)mat->beta;
  const int lnelas = mat->linear_elastic;
  const double alpha_zero = (double)mat->alpha_zero;
  const int nonlinear = iter >= 1 || extrapolated;
  double* ce = (double*)malloc(sizeof(double) * S * T);
  double* due = (double*)malloc(sizeof(double) * S * T);
  double(*dj)[MXGP] = malloc(sizeof(double) * S * MXGP);
  double(*gama)[MXGP][9] = malloc(sizeof(double) * S * MXGP * 9);
  double(*vol)[24] = malloc(sizeof(double) * S * 24);
  /* finite strains: displacements at n, [R n+1] and det F n+1 of every
     point, det J at n = 0 (rstgp1, gtmat1) */
  const int geonl = B->geonl;
  double* ue = geonl ? (double*)malloc(sizeof(double) * S * T) : NULL;
  double* rloc = geonl ? (double*)malloc(sizeof(double) * S * G * 9) : NULL;
  double* dfn1 = geonl ? (double*)malloc(sizeof(double) * S * G) : NULL;
  double* dj0 = geonl ? (double*)malloc(sizeof(double) * S * G) : NULL;
  double* jlast = geonl ? (double*)malloc(sizeof(double) * S * 9) : NULL;
  /* local n-state copies (gathered, then possibly reset at step 1) */
  double* hloc = (double*)malloc(sizeof(double) * S * G * H);
  double* sloc = (double*)malloc(sizeof(double) * S * G * 9);
  memcpy(hloc, B->hist_n, sizeof(double) * S * G * H);
  memcpy(sloc, B->urcs_n, sizeof(double) * S * G * 9);
  for (int i = 0; i < S; ++i) {
    gather_coords(om, B, i, ce + i * T);
    const int32_t* ed = B->edest + (int64_t)i * T;
    const double* tm[MXND];
    const int trn_e = elem_trn(om, B, i, tm);
    for (int k = 0; k < T; ++k) due[i * T + k] = du[ed[k] - 1];
    /* dupstr: displacements of skewed nodes to global coordinates */
    if (trn_e) trnvec(due + i * T, N, tm, 2);
    if (geonl) {
      /* rknstr_geonl: Jacobians and B-bar volume terms at the mid-step
         configuration ce_0 + u + du/2 (dupstr_blocked) */
      double cm[MXDOF];
      for (int k = 0; k < T; ++k) ue[i * T + k] = u[ed[k] - 1];
      if (trn_e) trnvec(ue + i * T, N, tm, 2);
      for (int k = 0; k < T; ++k)
        cm[k] = ce[i * T + k] + ue[i * T + k] + 0.5 * due[i * T + k];
      elem_kinematics(B, &t, cm, dj[i], gama[i], vol[i]);
    } else {
      elem_kinematics(B, &t, ce + i * T, dj[i], gama[i], vol[i]);
    }
  }
  double blk_energy = 0.0, blk_pwork = 0.0;
  double bmat[6 * MXDOF];
  int rc = 0, cut = 0;
  /* mm01_set_history at step 1 resets the gathered n-history of every
     element before any element is updated: the first element's word 11
     read by mm01_general (mm01.f:845) is then zero. */
  for (int g = 0; g < G; ++g) {
    double first_hist11 = hloc[((int64_t)0 * G + g) * H + 10];
    for (int i = 0; i < S; ++i) {
      int64_t q = (int64_t)i * G + g;
      /* gtlsn1: B, B-bar, deps = B due */
      orc_bmatrix(N, gama[i][g], t.nxi[g], t.neta[g], t.nzeta[g],
                  B->bbar ? vol[i] : NULL, om->eps_bbar, bmat);
      const double* d = due + i * T;
      double deps[6] = {0, 0, 0, 0, 0, 0};
      double qnhalf[36];
      if (geonl) {
        /* gtmat1: F at n+1 and n+1/2 from the n = 0 Jacobian, polar
           decompositions, [R n+1] kept, qnhalf from [R n+1/2] */
        double jac0[9], g0[9], th[9], fm[9], df, rn[9];
        orc_jacob1(N, t.nxi[g], t.neta[g], t.nzeta[g], ce + i * T, jac0,
                   &dj0[q], g0);
        /* the Jacobian left in local_work%jac after the last point */
        if (g == G - 1) memcpy(jlast + 9 * i, jac0, sizeof jac0);
        double un[MXDOF];
        for (int k = 0; k < T; ++k) un[k] = ue[i * T + k] + d[k];
        orc_tcomp1(N, g0, t.nxi[g], t.neta[g], t.nzeta[g], un, th);
        if (orc_fcomp1(th, fm, &dfn1[q])) cut = 2;
        orc_rtcmp1(fm, rloc + 9 * q);
        for (int k = 0; k < T; ++k) un[k] = ue[i * T + k] + 0.5 * d[k];
        orc_tcomp1(N, g0, t.nxi[g], t.neta[g], t.nzeta[g], un, th);
        if (orc_fcomp1(th, fm, &df)) cut = 2;
        orc_rtcmp1(fm, rn);
        orc_getrm1(rn, 1, qnhalf);
      }
      for (int j = 0; j < N; ++j) {
        int j1 = j, j2 = N + j, j3 = 2 * N + j;
        deps[0] = deps[0] + bmat[0 * T + j1] * d[j1] + bmat[0 * T + j2] * d[j2] +
                  bmat[0 * T + j3] * d[j3];
        deps[1] = deps[1] + bmat[1 * T + j1] * d[j1] + bmat[1 * T + j2] * d[j2] +
                  bmat[1 * T + j3] * d[j3];
        deps[2] = deps[2] + bmat[2 * T + j1] * d[j1] + bmat[2 * T + j2] * d[j2] +
                  bmat[2 * T + j3] * d[j3];
        deps[3] = deps[3] + bmat[3 * T + j1] * d[j1] + bmat[3 * T + j2] * d[j2];
        deps[4] = deps[4] + bmat[4 * T + j2] * d[j2] + bmat[4 * T + j3] * d[j3];
        deps[5] = deps[5] + bmat[5 * T + j1] * d[j1] + bmat[5 * T + j3] * d[j3];
      }
      if (geonl) { /* uddt = [qnhalf] ddt (qmply1) */
        double ddt[6];
        memcpy(ddt, deps, sizeof ddt);
        orc_qmply1(qnhalf, ddt, deps);
      }
      /* rstgp1_update_strains */
      for (int k = 0; k < 6; ++k)
        B->eps_n1[q * 6 + k] = B->eps_n[q * 6 + k] + deps[k];
      /* gauss_pt_temps, the point's properties, and the thermal strain
         increment taken from the increment the material sees
         (drive_01_update rstgp1.f:770-819, drive_10_update :4386-4447) */
      double dtn[MXND], tn1[MXND];
      elem_temps(om, B, i, dtn, tn1);
      const double gp_dtemps = om->temps_incr ? gp_interp(N, t.sf[g], dtn)
                                              : 0.0;
      const double gp_temps = gp_interp(N, t.sf[g], tn1);
      pt_props pp;
      point_props(mat, gp_temps, gp_dtemps, &pp);
      if (om->temps_incr)
        thermal_strain(pp.alpha, pp.alpha_n, alpha_zero, gp_temps,
                       gp_dtemps, deps);
      const double e = pp.e, nu = pp.nu;
      double* cgn = sloc + q * 9;
      double* cgn1 = B->urcs_n1 + q * 9;
      double cep[36];
      if (B->model == W3D_MM10) {
        /* drive_10_update: rotation I (small strain); iteration 0 is the
           linear-elastic update unless the increment was extrapolated
           after step 1 (rstgp1.f:4453-4457); history always scattered */
        int iter0_off = iter == 0 && (!extrapolated || step == 1);
        double* h1 = B->hist_n1 + q * H;
        /* finite strains: [R n+1] of the point (rot_blk_n1 layout) */
        double rcm[9];
        if (geonl)
          for (int a = 0; a < 3; ++a)
            for (int c2 = 0; c2 < 3; ++c2)
              rcm[a + 3 * c2] = rloc[9 * q + 3 * a + c2];
        if (orc_mm10_gp(mat, B->crys + (int64_t)i * B->ncry, step, iter0_off,
                        om->dt, deps, cgn, cgn1, hloc + q * H, h1,
                        om->asym, geonl ? rcm : NULL))
          cut = 1;
        /* [D] as stored for the tangent: history words 1-36 (6x6,
           column-major, drive_10_cnst) */
        for (int r = 0; r < 6; ++r)
          for (int cc = 0; cc < 6; ++cc) cep[6 * r + cc] = h1[6 * cc + r];
      } else if (nonlinear) {
        if (step == 1) first_hist11 = 0.0;
        double h1[11];
        if (orc_mm01_gp_t(step, e, nu, pp.e_n, pp.nu_n, beta, pp.hp, lnelas,
                          pp.yld, cgn, cgn1, deps, hloc + q * H, h1,
                          first_hist11, cep) != 0)
          rc = W3D_ERR_NUMERIC;
        /* rplstr: history is scattered for iter > 0 only */
        if (iter > 0) memcpy(B->hist_n1 + q * H, h1, sizeof h1);
      } else {
        /* drive_01_update_a / drive_01_update_b: linear-elastic [D] and
           sigma_n+1(1:6) = sigma_n + [D] deps; words 7-9 stay zero */
        const double one = 1.0, two = 2.0;
        double c1 = (e / ((one + nu) * (one - two * nu)));
        double c2 = (one - nu) * c1;
        double c3 = ((one - two * nu) / two) * c1;
        double c4 = nu * c1;
        memset(cep, 0, sizeof cep);
        cep[0] = c2; cep[7] = c2; cep[14] = c2;
        cep[21] = c3; cep[28] = c3; cep[35] = c3;
        cep[1] = c4; cep[2] = c4; cep[6] = c4; cep[12] = c4; cep[8] = c4;
        cep[13] = c4;
        for (int k = 0; k < 6; ++k) cgn1[k] = cgn[k];
        for (int k = 0; k < 6; ++k)
          for (int m = 0; m < 6; ++m)
            cgn1[k] = cgn1[k] + cep[6 * m + k] * deps[m];
        cgn1[6] = cgn1[7] = cgn1[8] = 0.0;
      }
      /* rstgp1_store_cep: 21-term lower triangle, layout (21, span, ngp) */
      double* cs = B->cep + ((int64_t)g * S + i) * 21;
      int k = 0;
      for (int r = 0; r < 6; ++r)
        for (int cc = 0; cc <= r; ++cc) cs[k++] = cep[6 * r + cc];
    }
    /* rstgp1_b, only for iter > 0: the deformed volume det F det J0 for
       finite strains */
    if (iter > 0) {
      for (int i = 0; i < S; ++i) {
        int64_t q = (int64_t)i * G + g;
        double factor = geonl ? dfn1[q] * dj0[q] * 1.0 * t.w[g]
                              : dj[i][g] * 1.0 * t.w[g];
        blk_energy = blk_energy + B->urcs_n1[q * 9 + 6] * factor;
        blk_pwork = blk_pwork + B->urcs_n1[q * 9 + 7] * factor;
      }
    }
  }
  /* rknstr_finish_cp: grad Fe (history words 37-63) of every element of
     a single-crystal block from the gathered n-history as mm10 left it
     (re-initialised in place at step 1); zero with several crystals */
  if (B->model == W3D_MM10 && !cut && !geonl) {
    double gv[27];
    if (B->ncry == 1)
      orc_mm10_calc_grads(B->etype, B->order, G, S, H, hloc, gv);
    else
      memset(gv, 0, sizeof gv);
    for (int64_t q = 0; q < (int64_t)S * G; ++q)
      memcpy(B->hist_n1 + q * H + 36, gv, sizeof gv);
  }
  /* finite strains: per element, Rp R^T fitted and taken to the current
     axes with the element's last n = 0 Jacobian (mm10_calc_grads) */
  if (B->model == W3D_MM10 && !cut && geonl)
    for (int i = 0; i < S; ++i) {
      double gv[27], rcm[9 * MXGP];
      memset(gv, 0, sizeof gv);
      if (B->ncry == 1) {
        for (int g = 0; g < G; ++g)
          for (int a = 0; a < 3; ++a)
            for (int c2 = 0; c2 < 3; ++c2)
              rcm[9 * g + a + 3 * c2] =
                  rloc[9 * ((int64_t)i * G + g) + 3 * a + c2];
        orc_mm10_calc_grads_geonl(B->etype, B->order, G, H,
                                  hloc + (int64_t)i * G * H, rcm,
                                  jlast + 9 * i, gv);
      }
      for (int g = 0; g < G; ++g)
        memcpy(B->hist_n1 + ((int64_t)i * G + g) * H + 36, gv, sizeof gv);
    }
  /* rknifv_std: element internal forces with the updated stresses (the
     coordinates at n+1 equal the initial ones in the small-displacement
     form, so the Jacobians above are reused) */
  memset(B->einfvec, 0, sizeof(double) * S * T);
  memset(B->evol, 0, sizeof(double) * S);
  if (geonl) /* rknifv_std: kinematics at n+1 (ce_0 + u + du) */
    for (int i = 0; i < S; ++i) {
      double cn[MXDOF];
      for (int k = 0; k < T; ++k)
        cn[k] = ce[i * T + k] + ue[i * T + k] + due[i * T + k];
      elem_kinematics(B, &t, cn, dj[i], gama[i], vol[i]);
    }
  for (int g = 0; g < G; ++g) {
    double w = t.w[g] * scalar;
    for (int i = 0; i < S; ++i) {
      int64_t q = (int64_t)i * G + g;
      orc_bmatrix(N, gama[i][g], t.nxi[g], t.neta[g], t.nzeta[g],
                  B->bbar ? vol[i] : NULL, om->eps_bbar, bmat);
      const double* s = B->urcs_n1 + q * 9;
      double cs[6];
      if (geonl) { /* gpifv1: Cauchy stress {T} = [qn1] {t} at n+1 */
        double qn1[36];
        orc_getrm1(rloc + 9 * q, 2, qn1);
        orc_qmply1(qn1, s, cs);
        s = cs;
      }
      double d = dj[i][g];
      for (int j = 0; j < T; ++j) {
        double* f = &B->einfvec[(int64_t)j * S + i];
        *f = *f + bmat[0 * T + j] * s[0] * w * d + bmat[1 * T + j] * s[1] * w * d +
             bmat[2 * T + j] * s[2] * w * d + bmat[3 * T + j] * s[3] * w * d +
             bmat[4 * T + j] * s[4] * w * d + bmat[5 * T + j] * s[5] * w * d;
      }
      B->evol[i] = B->evol[i] + w * d;
    }
  }
  /* rknifv: element forces of skewed nodes to constraint-compatible
     coordinates (trnvecs, switch 1) */
  for (int i = 0; i < S; ++i) {
    const double* tm[MXND];
    if (!elem_trn(om, B, i, tm)) continue;
    double f[MXDOF];
    for (int k = 0; k < T; ++k) f[k] = B->einfvec[(int64_t)k * S + i];
    trnvec(f, N, tm, 1);
    for (int k = 0; k < T; ++k) B->einfvec[(int64_t)k * S + i] = f[k];
  }
  /* rplstr: the gathered history of iter <= 0 is not scattered back; the
     n+1 history array keeps its previous contents.  [R n+1] (column-major
     records) is scattered for iter > 0 only. */
  if (geonl && iter > 0)
    for (int64_t q = 0; q < (int64_t)S * G; ++q)
      for (int a = 0; a < 3; ++a)
        for (int c2 = 0; c2 < 3; ++c2)
          B->rot_n1[9 * q + a + 3 * c2] = rloc[9 * q + 3 * a + c2];
  free(ce); free(due); free(dj); free(gama); free(vol); free(hloc);
  free(sloc); free(ue); free(rloc); free(dfn1); free(dj0); free(jlast);
  if (cut == 2) /* det F <= 0: adaptive step reduction (rstgp1.f) */
    return step > 1 ? W3D_CUT_STEP : W3D_ERR_NUMERIC;
  *energy = blk_energy;
  *pwork = blk_pwork;
  B->energy = iter > 0 ? blk_energy : 0.0;
  B->pwork = iter > 0 ? blk_pwork : 0.0;
  return rc ? rc : (cut ? W3D_CUT_STEP : 0);
}

/* do_nleps_block's block_energy and block_plastic_work of block blk's
   last recovery (zero for iteration 0), and addifv_standard's sum of
   |eleifv(i,j)| over the block, j outer and i inner (addifv.f:100-105). */
int orc_block_sums(const orc_model* om, int blk, double* out3) {
  if (blk < 1 || blk > om->nelblk)
    return fail(W3D_ERR_ARG, "orc_block_sums: block");
  const orc_block* B = &om->blk[blk - 1];
  double s = 0.0;
  for (int j = 0; j < B->totdof; ++j)
    for (int i = 0; i < B->span; ++i)
      s = s + fabs(B->einfvec[(int64_t)j * B->span + i]);
  out3[0] = B->energy;
  out3[1] = B->pwork;
  out3[2] = s;
  return 0;
}

static void keep_displacements(orc_model* om, const double* u,
                               const double* du) {
  memcpy(om->u_last, u, sizeof(double) * 3 * (size_t)om->nonode);
  memcpy(om->du_last, du, sizeof(double) * 3 * (size_t)om->nonode);
}

int orc_recover_blocks(orc_model* om, int first, int n, int step, int iter,
                       int extrapolated, const double* u, const double* du) {
  if (first < 1 || n < 0 || first - 1 + n > om->nelblk)
    return fail(W3D_ERR_ARG, "orc_recover_blocks: block range");
  keep_displacements(om, u, du);
  int rc = 0, cut = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc) \
    reduction(max : cut) \
    num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int b = first - 1; b < first - 1 + n; ++b) {
    double en, pw;
    int r = recover_block(om, b, step, iter, extrapolated, u, du, &en, &pw);
    if (r == W3D_CUT_STEP) cut = 1;
    else if (r < rc) rc = r;
  }
  if (rc) return fail(rc, "mm01: negative plastic multiplier");
  if (cut) return fail(W3D_CUT_STEP, "mm10: local solve failed, cut step");
  return 0;
}

int orc_recover(orc_model* om, int step, int iter, int extrapolated,
                const double* u, const double* du, double* ifv,
                double* energy, double* plastic_work) {
  int nb = om->nelblk;
  keep_displacements(om, u, du);
  double* be = (double*)calloc(nb, sizeof(double));
  double* bp = (double*)calloc(nb, sizeof(double));
  int rc = 0, cut = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : rc) \
    reduction(max : cut) \
    num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int b = 0; b < nb; ++b) {
    int r = recover_block(om, b, step, iter, extrapolated, u, du, &be[b],
                          &bp[b]);
    if (r == W3D_CUT_STEP) cut = 1;
    else if (r < rc) rc = r;
  }
  double en = 0.0, pw = 0.0;
  for (int b = 0; b < nb; ++b) {
    en = en + be[b];
    pw = pw + bp[b];
  }
  free(be);
  free(bp);
  if (energy) *energy = en;
  if (plastic_work) *plastic_work = pw;
  if (rc) return fail(rc, "mm01: negative plastic multiplier");
  /* assembly of element forces: per structure dof, over the node's
     elements in ascending order */
  if (ifv) {
#pragma omp parallel for schedule(static) \
    num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
    for (int n = 0; n < om->nonode; ++n) {
      double f[3] = {0, 0, 0};
      for (int p = om->inv_ptr[n]; p < om->inv_ptr[n + 1]; ++p) {
        int e = om->inv_list[p] - 1;
        const orc_block* B = &om->blk[om->elem_blk[e]];
        int i = e - (B->felem - 1);
        const int32_t* ed = B->edest + (int64_t)i * B->totdof;
        for (int j = 0; j < B->totdof; ++j) {
          int d = ed[j] - 1;
          if (d / 3 == n) f[d % 3] += B->einfvec[(int64_t)j * B->span + i];
        }
      }
      ifv[3 * n] = f[0];
      ifv[3 * n + 1] = f[1];
      ifv[3 * n + 2] = f[2];
    }
  }
  if (cut) return fail(W3D_CUT_STEP, "mm10: local solve failed, cut step");
  return 0;
}

/* ------------------------------------------------------------------ */
/* tangent: do_nlek_block / rktstf / gptns1 / bdbtgen                  */
/* ------------------------------------------------------------------ */
static void tangent_block(orc_model* om, int b) {
  orc_block* B = &om->blk[b];
  gp_table t;
  make_table(B, &t);
  const int S = B->span, G = B->ngp, N = B->nnode, T = B->totdof,
            U = B->utsz;
  const double scalar = B->etype == W3D_TET10 ? 1.0 / 6.0 : 1.0;
  double ce[MXDOF], dj[MXGP], gama[MXGP][9], vol[24];
  double bmat[6 * MXDOF], bd[6 * MXDOF];
  /* icp table, setup.f:268-276 */
  int* icp_r = (int*)malloc(sizeof(int) * U);
  int* icp_c = (int*)malloc(sizeof(int) * U);
  for (int j = 1; j <= T; ++j)
    for (int i = 1; i <= j; ++i) {
      int pos = (j - 1) * j / 2 + i;
      icp_r[pos - 1] = i;
      icp_c[pos - 1] = j;
    }
  for (int i = 0; i < S; ++i) {
    gather_coords(om, B, i, ce);
    const double* tm[MXND];
    const int trn_e = elem_trn(om, B, i, tm);
    if (B->geonl) { /* dptstf: coordinates updated to ce_0 + u + du, the
                       displacements of skewed nodes rotated to global */
      const int32_t* ed = B->edest + (int64_t)i * T;
      double ue[MXDOF], due[MXDOF];
      for (int k = 0; k < T; ++k) {
        ue[k] = om->u_last[ed[k] - 1];
        due[k] = om->du_last[ed[k] - 1];
      }
      if (trn_e) {
        trnvec(ue, N, tm, 2);
        trnvec(due, N, tm, 2);
      }
      for (int k = 0; k < T; ++k) ce[k] = ce[k] + ue[k] + due[k];
    }
    elem_kinematics(B, &t, ce, dj, gama, vol);
    double* ek = B->ek + (int64_t)i * U; /* transposed layout (utsz,span) */
    memset(ek, 0, sizeof(double) * U);
    double* ef = om->asym ? B->ekf + (int64_t)i * T * T : NULL;
    if (ef) memset(ef, 0, sizeof(double) * T * T);
    for (int g = 0; g < G; ++g) {
      orc_bmatrix(N, gama[g], t.nxi[g], t.neta[g], t.nzeta[g],
                  B->bbar ? vol : NULL, om->eps_bbar, bmat);
      double f = t.w[g] * dj[g];
      double dm[36];
      double qn1[36], cs[6];
      if (B->geonl) { /* gptns1_a: [qn1] from the stored [R n+1] */
        const double* rc = B->rot_n1 + 9 * ((int64_t)i * G + g);
        double r9[9];
        for (int a = 0; a < 3; ++a)
          for (int c2 = 0; c2 < 3; ++c2) r9[3 * a + c2] = rc[a + 3 * c2];
        orc_getrm1(r9, 2, qn1);
        orc_qmply1(qn1, B->urcs_n1 + 9 * ((int64_t)i * G + g), cs);
      }
      if (B->model == W3D_MM10) {
        /* drive_10_cnst (gptns1.f:1371): d(m,k) = history1 word 6k+m
           times w |J| */
        const double* h1 = B->hist_n1 + ((int64_t)i * G + g) * B->hist;
        for (int m = 0; m < 6; ++m)
          for (int kk = 0; kk < 6; ++kk) dm[6 * m + kk] = h1[6 * kk + m] * f;
      } else {
        /* extract_D_symmetric: f = w |J| times the stored lower triangle */
        const double* cs = B->cep + ((int64_t)g * S + i) * 21;
        int k = 0;
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c <= r; ++c) {
            double v = f * cs[k++];
            dm[6 * r + c] = v;
            dm[6 * c + r] = v;
          }
      }
      if (B->geonl) /* ctran1 with the [Q-bar] correction (qbar_flag) */
        orc_ctran1(dm, qn1, cs, 1, dj[g] * t.w[g]);
      if (scalar != 1.0) /* adjust_cnst */
        for (int q = 0; q < 36; ++q) dm[q] = scalar * dm[q];
      /* bdbtgen: bd(j,k) = sum_m d(m,k) b(j,m) */
      for (int j = 0; j < T; ++j)
        for (int kk = 0; kk < 6; ++kk)
          bd[kk * T + j] = dm[0 * 6 + kk] * bmat[0 * T + j] +
                           dm[1 * 6 + kk] * bmat[1 * T + j] +
                           dm[2 * 6 + kk] * bmat[2 * T + j] +
                           dm[3 * 6 + kk] * bmat[3 * T + j] +
                           dm[4 * 6 + kk] * bmat[4 * T + j] +
                           dm[5 * 6 + kk] * bmat[5 * T + j];
      if (ef) { /* bdbt_asym: word j = col T + row */
        int j = 0;
        for (int col = 0; col < T; ++col)
          for (int row = 0; row < T; ++row, ++j)
            ef[j] = ef[j] + bmat[0 * T + col] * bd[0 * T + row] +
                    bmat[1 * T + col] * bd[1 * T + row] +
                    bmat[2 * T + col] * bd[2 * T + row] +
                    bmat[3 * T + col] * bd[3 * T + row] +
                    bmat[4 * T + col] * bd[4 * T + row] +
                    bmat[5 * T + col] * bd[5 * T + row];
        continue;
      }
      for (int j = 0; j < U; ++j) {
        int row = icp_r[j] - 1, col = icp_c[j] - 1;
        ek[j] = ek[j] + bmat[0 * T + col] * bd[0 * T + row] +
                bmat[1 * T + col] * bd[1 * T + row] +
                bmat[2 * T + col] * bd[2 * T + row] +
                bmat[3 * T + col] * bd[3 * T + row] +
                bmat[4 * T + col] * bd[4 * T + row] +
                bmat[5 * T + col] * bd[5 * T + row];
      }
      if (B->geonl) /* kgstiff: initial-stress stiffness (no 1/6 for tets) */
        orc_kgstiff(N, gama[g], t.nxi[g], t.neta[g], t.nzeta[g],
                    B->bbar ? vol : NULL, cs, dj[g], t.w[g], ek);
    }
    if (trn_e) { /* rktstf: trnmtx on the element's matrix */
      double full[MXDOF * MXDOF];
      if (ef) {
        memcpy(full, ef, sizeof(double) * T * T);
      } else {
        for (int j = 1; j <= T; ++j)
          for (int r = 1; r <= j; ++r) {
            double v = ek[(j - 1) * j / 2 + r - 1];
            full[(r - 1) + T * (j - 1)] = v;
            full[(j - 1) + T * (r - 1)] = v;
          }
      }
      trnmtx(full, N, tm);
      if (ef) {
        memcpy(ef, full, sizeof(double) * T * T);
      } else {
        for (int j = 1; j <= T; ++j)
          for (int r = 1; r <= j; ++r)
            ek[(j - 1) * j / 2 + r - 1] = full[(r - 1) + T * (j - 1)];
      }
    }
  }
  free(icp_r);
  free(icp_c);
}

int orc_tangent_blocks(orc_model* om, int first, int n, int step, int iter) {
  (void)step;
  (void)iter;
  if (first < 1 || n < 0 || first - 1 + n > om->nelblk)
    return fail(W3D_ERR_ARG, "orc_tangent_blocks: block range");
  for (int b = first - 1; b < first - 1 + n && om->asym; ++b) {
    orc_block* B = &om->blk[b];
    if (B->geonl)
      return fail(W3D_ERR_UNSUPPORTED,
                  "asymmetric assembly of finite-strain blocks");
    if (!B->ekf)
      B->ekf = (double*)malloc(sizeof(double) * B->totdof * B->totdof *
                               (size_t)B->span);
  }
#pragma omp parallel for schedule(dynamic, 1) \
    num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int b = first - 1; b < first - 1 + n; ++b) tangent_block(om, b);
  return 0;
}

int orc_tangent(orc_model* om, int step, int iter) {
  return orc_tangent_blocks(om, 1, om->nelblk, step, iter);
}

/* ------------------------------------------------------------------ */
/* sparsity and assembly                                               */
/* ------------------------------------------------------------------ */
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

int orc_pattern(orc_model* om, const int32_t* cstmap, int32_t* neqns_out,
                int64_t* ncoeff_out) {
  int nd = 3 * om->nonode;
  free(om->dof_eqn_map); free(om->eqn_node_map); free(om->k_ptrs);
  free(om->k_indexes); free(om->row_start);
  om->dof_eqn_map = (int32_t*)calloc(nd, sizeof(int32_t));
  om->eqn_node_map = (int32_t*)calloc(nd, sizeof(int32_t));
  /* dof_map, assemble_code.f:561 */
  int neq = 0;
  for (int d = 1; d <= nd; ++d) {
    if (cstmap[d - 1] != 0) continue;
    neq = neq + 1;
    om->dof_eqn_map[d - 1] = neq;
    om->eqn_node_map[neq - 1] = (d - 1) / 3 + 1;
  }
  om->neqns = neq;
  om->k_ptrs = (int32_t*)calloc(neq > 0 ? neq : 1, sizeof(int32_t));
  /* build_col_sparse_symm_srow, per row: collect, sort, unique */
  int32_t** rows = (int32_t**)calloc(neq > 0 ? neq : 1, sizeof(int32_t*));
#pragma omp parallel for schedule(dynamic, 256) \
    num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int srow = 1; srow <= neq - 1; ++srow) {
    int snode = om->eqn_node_map[srow - 1];
    int cap = 0;
    for (int p = om->inv_ptr[snode - 1]; p < om->inv_ptr[snode]; ++p) {
      int e = om->inv_list[p] - 1;
      cap += 3 * orc_nnode(om->etype[e]);
    }
    int32_t* list = (int32_t*)malloc(sizeof(int32_t) * (cap + 1));
    int nused = 0;
    for (int p = om->inv_ptr[snode - 1]; p < om->inv_ptr[snode]; ++p) {
      int e = om->inv_list[p] - 1;
      const orc_block* B = &om->blk[om->elem_blk[e]];
      int i = e - (B->felem - 1);
      const int32_t* ed = B->edest + (int64_t)i * B->totdof;
      for (int erow = 0; erow < B->totdof; ++erow) {
        if (om->dof_eqn_map[ed[erow] - 1] != srow) continue;
        for (int ecol = 0; ecol < B->totdof; ++ecol) {
          int scol = om->dof_eqn_map[ed[ecol] - 1];
          if (scol <= srow) continue;
          list[nused++] = scol;
        }
        break; /* no repeated incidences in these meshes */
      }
    }
    if (nused > 1) qsort(list, nused, sizeof(int32_t), cmp_i32);
    int nu = 0;
    for (int k = 0; k < nused; ++k)
      if (k == 0 || list[k] != list[k - 1]) list[nu++] = list[k];
    rows[srow - 1] = list;
    om->k_ptrs[srow - 1] = nu;
  }
  if (neq > 0) om->k_ptrs[neq - 1] = 0;
  om->row_start = (int64_t*)malloc(sizeof(int64_t) * (neq + 1));
  om->row_start[0] = 0;
  for (int r = 0; r < neq; ++r)
    om->row_start[r + 1] = om->row_start[r] + om->k_ptrs[r];
  om->ncoeff = om->row_start[neq];
  om->k_indexes = (int32_t*)malloc(sizeof(int32_t) * (om->ncoeff + 1));
  for (int r = 0; r < neq - 1; ++r) {
    if (om->k_ptrs[r])
      memcpy(om->k_indexes + om->row_start[r], rows[r],
             sizeof(int32_t) * om->k_ptrs[r]);
    free(rows[r]);
  }
  free(rows);
  if (neqns_out) *neqns_out = neq;
  if (ncoeff_out) *ncoeff_out = om->ncoeff;
  return 0;
}

int orc_get_pattern(const orc_model* om, int32_t* dof_eqn_map,
                    int32_t* eqn_node_map, int32_t* k_ptrs,
                    int32_t* k_indexes) {
  if (!om->dof_eqn_map) return fail(W3D_ERR_ARG, "no pattern");
  if (dof_eqn_map)
    memcpy(dof_eqn_map, om->dof_eqn_map, sizeof(int32_t) * 3 * om->nonode);
  if (eqn_node_map)
    memcpy(eqn_node_map, om->eqn_node_map, sizeof(int32_t) * om->neqns);
  if (k_ptrs) memcpy(k_ptrs, om->k_ptrs, sizeof(int32_t) * om->neqns);
  if (k_indexes)
    memcpy(k_indexes, om->k_indexes, sizeof(int32_t) * om->ncoeff);
  return 0;
}

/* assem_by_row / assem_a_row, assemble_code.f:659,785 */
int orc_assemble_vss(orc_model* om, double* k_diag, double* k_coeffs) {
  if (!om->dof_eqn_map) return fail(W3D_ERR_ARG, "no pattern");
  int neq = om->neqns;
#pragma omp parallel num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  {
    double* coeff_row = (double*)calloc(neq + 1, sizeof(double));
#pragma omp for schedule(dynamic, 256)
    for (int srow = 1; srow <= neq; ++srow) {
      int snode = om->eqn_node_map[srow - 1];
      for (int p = om->inv_ptr[snode - 1]; p < om->inv_ptr[snode]; ++p) {
        int e = om->inv_list[p] - 1;
        const orc_block* B = &om->blk[om->elem_blk[e]];
        int i = e - (B->felem - 1);
        const int32_t* ed = B->edest + (int64_t)i * B->totdof;
        const double* emat = B->ek + (int64_t)i * B->utsz;
        int T = B->totdof;
        int lsc[MXDOF];
        for (int k = 0; k < T; ++k) lsc[k] = om->dof_eqn_map[ed[k] - 1];
        for (int erow = 1; erow <= T; ++erow) {
          if (lsc[erow - 1] != srow) continue;
          for (int ecol = 1; ecol <= T; ++ecol) {
            int scol = lsc[ecol - 1];
            if (scol < srow) continue;
            int mx = ecol > erow ? ecol : erow;
            int dcp = (mx - 1) * mx / 2 + mx;
            int kk = dcp - abs(ecol - erow);
            coeff_row[scol] = coeff_row[scol] + emat[kk - 1];
          }
          break;
        }
      }
      k_diag[srow - 1] = coeff_row[srow];
      coeff_row[srow] = 0.0;
      if (srow == neq) continue;
      int64_t st = om->row_start[srow - 1];
      for (int j = 0; j < om->k_ptrs[srow - 1]; ++j) {
        int col = om->k_indexes[st + j];
        k_coeffs[st + j] = coeff_row[col];
        coeff_row[col] = 0.0;
      }
    }
    free(coeff_row);
  }
  return 0;
}

int orc_set_asymmetric(orc_model* om, int on) {
  om->asym = on != 0;
  return 0;
}

/* assem_by_row_asymmetric / assem_a_row_asym (assemble_code.f:1050): the
   row of equation srow gathers, over the elements on its node in
   inverse-incidence order, the Ke words (erow - 1) T + ecol of the
   element row erow that maps to srow */
int orc_assemble_csr(orc_model* om, const int64_t* rp, const int32_t* col,
                     double* val) {
  if (!om->dof_eqn_map) return fail(W3D_ERR_ARG, "no pattern");
  if (!om->asym) return fail(W3D_ERR_ARG, "asymmetric assembly is off");
  for (int b = 0; b < om->nelblk; ++b)
    if (!om->blk[b].ekf) return fail(W3D_ERR_ARG, "no full Ke: run tangent");
  int neq = om->neqns, bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : bad) \
    num_threads(g_threads > 0 ? g_threads : omp_get_max_threads())
  for (int srow = 1; srow <= neq; ++srow) {
    const int64_t r0 = rp[srow - 1], r1 = rp[srow];
    for (int64_t k = r0; k < r1; ++k) val[k] = 0.0;
    int snode = om->eqn_node_map[srow - 1];
    for (int p = om->inv_ptr[snode - 1]; p < om->inv_ptr[snode]; ++p) {
      int e = om->inv_list[p] - 1;
      const orc_block* B = &om->blk[om->elem_blk[e]];
      int i = e - (B->felem - 1);
      const int32_t* ed = B->edest + (int64_t)i * B->totdof;
      int T = B->totdof;
      const double* emat = B->ekf + (int64_t)i * T * T;
      int lsc[MXDOF];
      for (int k = 0; k < T; ++k) lsc[k] = om->dof_eqn_map[ed[k] - 1];
      for (int erow = 0; erow < T; ++erow) {
        if (lsc[erow] != srow) continue;
        for (int ecol = 0; ecol < T; ++ecol) {
          int scol = lsc[ecol];
          if (scol == 0) continue;
          int64_t lo = r0, hi = r1 - 1;
          while (lo < hi) {
            int64_t mid = (lo + hi) / 2;
            if (col[mid] < scol - 1) lo = mid + 1; else hi = mid;
          }
          if (lo >= r1 || col[lo] != scol - 1) { bad = 1; continue; }
          val[lo] = val[lo] + emat[erow * T + ecol];
        }
        break; /* no repeated incidences in these meshes */
      }
    }
  }
  return bad ? fail(W3D_ERR_ARG, "orc_assemble_csr: pattern lacks a column")
             : 0;
}

int orc_get_ek_full(const orc_model* om, int blk, double* out,
                    int64_t n_out) {
  if (blk < 1 || blk > om->nelblk) return fail(W3D_ERR_ARG, "block");
  const orc_block* B = &om->blk[blk - 1];
  int64_t n = (int64_t)B->totdof * B->totdof * B->span;
  if (!B->ekf) return fail(W3D_ERR_ARG, "no full Ke: run tangent");
  if (n_out < n) return fail(W3D_ERR_ARG, "output too small");
  memcpy(out, B->ekf, sizeof(double) * n);
  return 0;
}

/* ------------------------------------------------------------------ */
/* state roll-over and host views                                      */
/* ------------------------------------------------------------------ */
int orc_commit_step(orc_model* om) {
  for (int b = 0; b < om->nelblk; ++b) {
    orc_block* B = &om->blk[b];
    int64_t nq = (int64_t)B->span * B->ngp;
    memcpy(B->urcs_n, B->urcs_n1, sizeof(double) * 9 * nq);
    memcpy(B->eps_n, B->eps_n1, sizeof(double) * 6 * nq);
    memcpy(B->hist_n, B->hist_n1, sizeof(double) * B->hist * nq);
    if (B->geonl) memcpy(B->rot_n, B->rot_n1, sizeof(double) * 9 * nq);
  }
  return 0;
}

int orc_rollback(orc_model* om) {
  for (int b = 0; b < om->nelblk; ++b) {
    orc_block* B = &om->blk[b];
    int64_t nq = (int64_t)B->span * B->ngp;
    memcpy(B->urcs_n1, B->urcs_n, sizeof(double) * 9 * nq);
    memcpy(B->eps_n1, B->eps_n, sizeof(double) * 6 * nq);
    memcpy(B->hist_n1, B->hist_n, sizeof(double) * B->hist * nq);
    if (B->geonl) memcpy(B->rot_n1, B->rot_n, sizeof(double) * 9 * nq);
  }
  return 0;
}

static double* state_ptr(const orc_block* B, int which, int64_t* n) {
  int64_t nq = (int64_t)B->span * B->ngp;
  switch (which) {
    case W3D_STRESS_N: *n = 9 * nq; return B->urcs_n;
    case W3D_STRESS_N1: *n = 9 * nq; return B->urcs_n1;
    case W3D_STRAIN_N: *n = 6 * nq; return B->eps_n;
    case W3D_STRAIN_N1: *n = 6 * nq; return B->eps_n1;
    case W3D_HIST_N: *n = B->hist * nq; return B->hist_n;
    case W3D_HIST_N1: *n = B->hist * nq; return B->hist_n1;
    case W3D_CEP: *n = 21 * nq; return B->cep;
    case W3D_ESTIFF: *n = (int64_t)B->utsz * B->span; return B->ek;
    case W3D_EINFVEC: *n = (int64_t)B->totdof * B->span; return B->einfvec;
    case W3D_ELEM_VOL: *n = B->span; return B->evol;
    case W3D_ROT_N1: *n = B->geonl ? 9 * nq : 0; return B->rot_n1;
    default: *n = 0; return NULL;
  }
}

int orc_get_block_state(const orc_model* om, int blk, int which, double* out,
                        int64_t n_out) {
  if (blk < 1 || blk > om->nelblk) return fail(W3D_ERR_ARG, "bad block");
  int64_t n;
  double* p = state_ptr(&om->blk[blk - 1], which, &n);
  if (!p || n_out < n) return fail(W3D_ERR_ARG, "bad selector or size");
  memcpy(out, p, sizeof(double) * n);
  return 0;
}

int orc_set_block_state(orc_model* om, int blk, int which, const double* in,
                        int64_t n_in) {
  if (blk < 1 || blk > om->nelblk) return fail(W3D_ERR_ARG, "bad block");
  if (which != W3D_STRESS_N && which != W3D_STRAIN_N && which != W3D_HIST_N)
    return fail(W3D_ERR_ARG, "bad selector");
  int64_t n;
  double* p = state_ptr(&om->blk[blk - 1], which, &n);
  if (n_in < n) return fail(W3D_ERR_ARG, "bad size");
  memcpy(p, in, sizeof(double) * n);
  return 0;
}
