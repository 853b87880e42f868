/*
 * w3d_oracle.h -- CPU FP64 restatement of WARP3D's element-block hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker for libw3dblk: only
 * the tests, the smoke check and the CPU-baseline leg of bench.py load it.
 * The product path never calls it.
 *
 * Every function here restates a reference routine (file:line under the
 * reference's src/), in the reference's data layouts, so the tests can compare
 * libw3dblk's device results with it array by array.
 */
#ifndef W3D_ORACLE_H
#define W3D_ORACLE_H

#include <stdint.h>
#include "../include/w3d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_model orc_model;

/* ---- leaf routines (exported for the pinning tests) ------------------ */
int orc_ngp(int etype, int order);
int orc_nnode(int etype);
void orc_getgpts(int etype, int order, int gpn, double* xi, double* eta,
                 double* zeta, double* w);
void orc_derivs(int etype, double xi, double eta, double zeta, double* nxi,
                double* neta, double* nzeta);
/* jacob1 for one element: ce is (3N) x-block,y-block,z-block. */
void orc_jacob1(int nnode, const double* nxi, const double* neta,
                const double* nzeta, const double* ce, double jac[9],
                double* dj, double gama[9]);
/* B (6 x 3N, row-major [r][dof]) of blcmp1 (+ bmod when vol != NULL). */
void orc_bmatrix(int nnode, const double gama[9], const double* nxi,
                 const double* neta, const double* nzeta, const double* vol,
                 double eps_bbar, double* b);
/*
 * mm01 + cnst1 for one Gauss point of one element (mm01.f:24, :1196).
 * hist/hist1 are 11 words, cgn/cgn1 9 words, deps 6, cep 36 (6x6 row-major).
 * first_hist11 is history(1,11) of the block's first element at this point
 * (the quirk of mm01.f:845).  Returns 0, or -1 on a negative plastic
 * multiplier (die_abort in the reference).
 */
int orc_mm01_gp(int step, double e, double nu, double beta, double hprime,
                int lnelas, double yld, double* cgn, double* cgn1,
                const double* deps, double* hist, double* hist1,
                double first_hist11, double* cep);
/* The same with the moduli at n (e_n, nu_n) apart from those at n+1. */
int orc_mm01_gp_t(int step, double e, double nu, double e_n, double nu_n,
                  double beta, double hprime, int lnelas, double yld,
                  double* cgn, double* cgn1, const double* deps, double* hist,
                  double* hist1, double first_hist11, double* cep);
/* Shape-function values of an element type at a parametric point
   (shapef.f:15), node numbering as orc_derivs. */
void orc_shapef(int etype, double xi, double eta, double zeta, double* n);
int orc_set_temperatures(orc_model* om, const double* temper_nodes,
                         const double* dtemp_nodes, const double* temper_elems,
                         const double* dtemp_elems, int incremental);
/* H' = e*et/(e-et) in REAL arithmetic (elprp.f:872). */
float orc_hprime(float e, float et);

/* ---- mm10 (orc_mm10.c) ------------------------------------------------ */
/* One crystal of one element: stiffness rotated into the reference frame
   (i,j) at C[i + 6j], Schmid tensors ms (Voigt, engineering shear) and skew
   vectors qs of the 12 slip systems, the crystal -> reference rotation g
   ((i,j) at g[i + 3j]) and the initial Euler angles. */
typedef struct {
  double C[36];
  double ms[12][6];
  double qs[12][3];
  double g[9];
  double angles[3];
} orc_crystal;
int orc_mm10_crystal(const w3d_material* mat, const double angles[3],
                     orc_crystal* cr);
int orc_mm10_hist_size(int ncrystals);
/* mm10 for one Gauss point (all its crystals).  hn is the point's gathered
   n-history (re-initialised in place at step 1), hn1 receives the n+1
   history, urcs 9 words.  Returns 0, or 1 when the local solve fails (the
   reference's material_cut_step).  asym: the tangent is not symmetrised
   (asymmetric assembly, mm10_a.f:1165).  rn1: the point's [R n+1] from
   the polar decomposition (9 words, column-major like rot_blk_n1), NULL
   for small strains (identity, drive_10_update rstgp1.f:4405-4417). */
int orc_mm10_gp(const w3d_material* mat, const orc_crystal* crys, int step,
                int iter0_off, double dt, const double* uddt,
                const double* urcs_n, double* urcs_n1, double* hn,
                double* hn1, int asym, const double* rn1);
void orc_mm10_residual(const w3d_material* mat, const orc_crystal* cr,
                       const double* hist_crystal, const double* uddt,
                       double dt, const double* x, double* R, double* J,
                       const double* rn1);
void orc_mm10_rotation_matrix(const double angles[3], int angle_type,
                              double r[9]);
/* Work counters of the mm10 local solves since the last reset: [0] crystal
   updates, [1] sub-step solves, [2] formR1 evaluations (also those inside
   formR), [3] formR, [4] formJ11 (also inside formJ), [5] formJ, [6] 6x6 and
   [7] 7x7 LU solves; hist64: points by formR1 evaluations (bin 63: 63 or
   more).  Either output may be NULL. */
void orc_mm10_counters(long long* out8, long long* hist64, int reset);
/* finite strains (orc_geonl.c); 3x3 matrices row-major */
void orc_tcomp1(int nnode, const double gama[9], const double* nxi,
                const double* neta, const double* nzeta, const double* ue,
                double theta[9]);
int orc_fcomp1(const double theta[9], double f[9], double* df);
void orc_evcmp1(const double c[6], double ev[3]);
void orc_irscp1(const double f[9], double ui[6]);
void orc_rtcmp1(const double f[9], double r[9]);
void orc_getrm1(const double r9[9], int opt, double q[36]);
void orc_qmply1(const double q[36], const double m1[6], double m2[6]);
void orc_ctran1(double d[36], const double q[36], const double cs[6],
                int qbar, double wdj);
void orc_kgstiff(int nnode, const double gama[9], const double* nxi,
                 const double* neta, const double* nzeta, const double* vol,
                 const double sig[6], double dj, double w, double* ek);
/* Least squares min |A x - b|, A (m x 4) column-major (DGELS('N')). */
void orc_lsq4(int m, const double* a, const double* b, double* x);
/* grad Fe words (27) of single-crystal elements of a small-strain block
   (rknstr_finish_cp / mm10_calc_grads) from the block's gathered
   n-history hloc (hist, ngp, span). */
void orc_mm10_calc_grads(int etype, int order, int ngp, int span, int hist,
                         const double* hloc, double* grads_vec);
/* The same for one element of a finite-strain block (mm10_calc_grads'
   geonl branch): the fitted components are those of Rp R^T at each point
   (Rp of crystal 1 from the element's n-history hel (hist, ngp), R =
   [R n+1] from rot (9 column-major words a point)), and the slopes are
   taken to the current axes by the inverse of jac (the element's n = 0
   coordinate Jacobian at its last point, row-major).  Returns 1 when jac
   is singular (the reference leaves the inverse zero). */
int orc_mm10_calc_grads_geonl(int etype, int order, int ngp, int hist,
                              const double* hel, const double* rot,
                              const double jac[9], double* grads_vec);
void orc_mm10_rt2rve(const double* rt, double* rv);
void orc_mm10_rt2rvw(const double* rt, double* rv);
void orc_mm10_expmw(const double wv[3], double a[9]);

/* ---- block model ------------------------------------------------------ */
orc_model* orc_create(const w3d_model* m);
void orc_destroy(orc_model* om);
const char* orc_last_error(void);
int orc_get_blocking(const orc_model* om, int32_t* nelblk, int32_t* elblks);
int orc_block_info(const orc_model* om, int blk, int32_t* info5);
void orc_set_threads(int n);
void orc_set_time_increment(orc_model* om, double dt);

int orc_recover(orc_model* om, int step, int iter, int extrapolated,
                const double* u, const double* du, double* ifv,
                double* energy, double* plastic_work);
/* Recovery of blocks [first, first+n) only, element forces left in einfvec. */
int orc_block_sums(const orc_model* om, int blk, double* out3);
int orc_recover_blocks(orc_model* om, int first, int n, int step, int iter,
                       int extrapolated, const double* u, const double* du);
int orc_tangent(orc_model* om, int step, int iter);
int orc_tangent_blocks(orc_model* om, int first, int n, int step, int iter);
int orc_pattern(orc_model* om, const int32_t* cstmap, int32_t* neqns,
                int64_t* ncoeff);
int orc_get_pattern(const orc_model* om, int32_t* dof_eqn_map,
                    int32_t* eqn_node_map, int32_t* k_ptrs,
                    int32_t* k_indexes);
int orc_assemble_vss(orc_model* om, double* k_diag, double* k_coeffs);
/* Asymmetric assembly (the solution parameter asymmetric_assembly): mm10
   keeps its tangent unsymmetrised, the tangent forms the full Ke of every
   element (bdbt_asym, gptns1.f:1513) and orc_assemble_csr sums it into the
   values of a full CSR pattern given by the caller (rp: neqns+1 offsets,
   col: 0-based equations, ascending in each row) in the reference's order
   (assem_a_row_asym, assemble_code.f:1050). */
int orc_set_asymmetric(orc_model* om, int on);
int orc_assemble_csr(orc_model* om, const int64_t* rp, const int32_t* col,
                     double* val);
/* The full Ke of the elements of block blk (span x T*T, each in bdbt_asym's
   order: word col*T + row holds B(row)^T D B(col)). */
int orc_get_ek_full(const orc_model* om, int blk, double* out, int64_t n_out);
int orc_commit_step(orc_model* om);
int orc_rollback(orc_model* om);
int orc_get_block_state(const orc_model* om, int blk, int which, double* out,
                        int64_t n_out);
int orc_set_block_state(orc_model* om, int blk, int which, const double* in,
                        int64_t n_in);

#ifdef __cplusplus
}
#endif

#endif
