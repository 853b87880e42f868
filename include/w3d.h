/*
 * w3d.h -- C ABI of libw3dblk, the MI355X (gfx950) element-block library.
 *
 * The library replaces the element-block hot path of WARP3D (rhdodds/warp3d):
 * Gauss-point strain/stress recovery with the mm01 (bilinear Mises) and mm10
 * (crystal plasticity) material updates, the element internal-force vector,
 * the block tangent Ke = sum B^T D B w |J| and the assembly of the blocks into
 * the upper-triangular "NASA-VSS" sparse matrix.
 *
 * Conventions (all follow the reference's Fortran data):
 *   - every integer that names a node, element, block, dof or equation is
 *     1-based, exactly as in the reference arrays it mirrors;
 *   - every array is plain memory: host pointers in the w3d_* entry points,
 *     device pointers in the w3d_*_dev entry points (device pointers must come
 *     from hipMalloc on the context's device);
 *   - every entry point returns W3D_OK (0), W3D_CUT_STEP (1: a material or
 *     kinematic failure asks the caller to cut the load step) or a negative
 *     error code; w3d_last_error() gives the message of the last failure on
 *     the calling thread.
 *
 * Reference interfaces replaced (file:line under the reference's src/):
 *   w3d_create            model set-up in incomp.f:14 / setup.f / inelbk.f:282
 *   w3d_recover           drive_eps_sig_internal_forces.f:20
 *   w3d_recover_block     do_nleps_block, drive_eps_sig_internal_forces.f:399
 *   w3d_recover_blocks    the same for a range of blocks
 *   w3d_assemble_ifv      addifv.f:16 (the element forces into ifv)
 *   w3d_tangent           tanstf.f:17
 *   w3d_tangent_block     do_nlek_block, tanstf.f:119
 *   w3d_tangent_blocks    the same for a range of blocks
 *   w3d_pattern           drive_assemble_solve.f:358 (ds_setup_sparsity_local)
 *   w3d_assemble_vss      drive_assemble_solve.f:654 (ds_root_assembly ->
 *                         assemble_code.f:659 assem_by_row)
 *   w3d_commit_step       update.f:15 (n+1 -> n roll-over of block state)
 *   w3d_get_block_state   the block arrays of mod_eleblocks.f:17-161
 */
#ifndef W3D_H
#define W3D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define W3D_ABI_VERSION 5

/* curves of one mm01 segmental curve set (temperature-dependent form) */
#define W3D_MAX_CURVES 8

/* ---- status codes ---------------------------------------------------- */
enum {
  W3D_OK = 0,
  W3D_CUT_STEP = 1,          /* material_cut_step = .true. in the reference */
  W3D_ERR_ARG = -1,          /* invalid argument / inconsistent model       */
  W3D_ERR_UNSUPPORTED = -2,  /* feature outside the implemented scope       */
  W3D_ERR_DEVICE = -3,       /* HIP runtime failure                         */
  W3D_ERR_NOMEM = -4,        /* device or host allocation failure           */
  W3D_ERR_NUMERIC = -5       /* fatal numerical failure (reference die_abort) */
};

/* ---- element types: reference iprops(1,elem) (elprp.f) ---------------- */
enum {
  W3D_Q3DISOP = 1,  /* 20-node hexahedron                                 */
  W3D_L3DISOP = 2,  /* 8-node hexahedron                                  */
  W3D_TET10 = 6     /* 10-node tetrahedron                                */
};

/* ---- material models: reference iprops(25,elem) ---------------------- */
enum {
  W3D_MM01 = 1,     /* bilinear Mises, mixed iso/kinematic hardening      */
  W3D_MM10 = 10     /* crystal plasticity (FCC, Voce hardening)           */
};

/* ---- tangent precision ----------------------------------------------- */
enum {
  W3D_FP64 = 0,     /* B^T D B formed and stored in FP64                  */
  W3D_FP32 = 1      /* B^T D B formed on MFMA in FP32 with FP32 accumulate;
                       Ke stored in FP32 and widened during assembly      */
};

/* ---- block state selectors for w3d_get_block_state ------------------- */
enum {
  W3D_STRESS_N = 0,   /* urcs_n_blocks   (9, ngp, span)                   */
  W3D_STRESS_N1 = 1,  /* urcs_n1_blocks  (9, ngp, span)                   */
  W3D_STRAIN_N = 2,   /* eps_n_blocks    (6, ngp, span)                   */
  W3D_STRAIN_N1 = 3,  /* eps_n1_blocks   (6, ngp, span)                   */
  W3D_HIST_N = 4,     /* history_blocks  (hist, ngp, span)                */
  W3D_HIST_N1 = 5,    /* history1_blocks (hist, ngp, span)                */
  W3D_CEP = 6,        /* cep_blocks      (21, span, ngp)                  */
  W3D_ESTIFF = 7,     /* estiff_blocks   (utsz, span)                     */
  W3D_EINFVEC = 8,    /* einfvec_blocks  (span, totdof)                   */
  W3D_ELEM_VOL = 9,   /* element_vol_blocks (span)                        */
  W3D_ROT_N1 = 10     /* rot_n1_blocks   (9, ngp, span): [R n+1] of F = RU,
                         column-major; finite-strain blocks only         */
};

/*
 * One material.  The mm01 fields mirror the reference's REAL matprp rows
 * (mod_main.f:373): e (1), nu (2), beta (3), tan_e (4), yld_pt (5); they are
 * single precision on purpose, because the reference rounds them through
 * REAL and forms H' = e*tan_e/(e - tan_e) in REAL (elprp.f:722,872).
 * The mm10 fields are one crystal definition, the reference's double
 * precision c_array entry (mod_crystals.f:149-221, defaults at :238-417),
 * shared by every crystal of the material; the orientation of each crystal
 * comes from w3d_model.angles.
 */
typedef struct w3d_material {
  int32_t model;            /* W3D_MM01 or W3D_MM10                        */
  /* mm01 */
  float e, nu, beta, tan_e, yld_pt;
  int32_t linear_elastic;   /* lprops(17): response always linear elastic  */
  /* mm10 (Voce hardening); ignored for mm01 */
  int32_t slip_type;        /* 1 = fcc, 2 = bcc: 12 systems each           */
  int32_t elastic_type;     /* 1 = isotropic (cp_e, cp_nu),
                               2 = cubic (cp_e, cp_nu, cp_mu)              */
  double cp_e, cp_nu, cp_mu;/* crystal elastic constants                   */
  double harden_n;          /* slip-rate exponent (rate_n)                 */
  double theta_0;           /* Voce initial hardening rate                 */
  double tau_y, tau_v;      /* initial slip resistance, saturation gap     */
  double voce_m;            /* Voce exponent                               */
  double iD_v;              /* diffusion term                              */
  double eps_dot_0_y;       /* fixed rate prefactor when alter_mode = 1    */
  double atol, atol1, rtol, rtol1;  /* local Newton tolerances             */
  int32_t miter;            /* local Newton iteration limit                */
  int32_t alter_mode;       /* Voce alternative rate form                  */
  int32_t angle_type;       /* 1 = degrees, 2 = radians                    */
  int32_t angle_convention; /* 1 = Kocks (the only one the reference
                               accepts, rknstr.f:1927)                     */
  int32_t ncrystals;        /* crystals per point, Taylor-averaged         */
  int32_t pad_;
  /* thermal expansion (both models): props 9, 13, 34-37 of the element
     property table (REAL): alpha xx, yy, zz, xy, yz, xz (rknstr.f:1046-1053)
     and alpha_zero, matprp(250) (rknstr.f:150) */
  float alpha[6];
  float alpha_zero;
  /* mm01 segmental stress-strain curve set (the material's curve
     property; segmental_curves module, set_up_segmental,
     setup_mat_props.f:67): curve_type 0 none (H' from tan_e, yield
     yld_pt); 1 one temperature-independent curve, the reference's curve
     type 0 (curve 0 below; e, nu and alpha stay the properties above); 2
     temperature dependent, the reference's curve type 1: ncurves curves
     (at most W3D_MAX_CURVES) at the ascending
     temperatures curve_temp, each with its own e, nu and isotropic alpha,
     interpolated linearly at the point's temperatures at n and n+1
     (linear_interpolate, setup_mat_props.f:361, constant outside the
     range).  mm01 uses the first two points of the curve at the point's
     temperature (set_up_h_prime, setup_mat_props.f:406): H' = (sig_2 -
     sig_1) / (eps_2 - eps_1), yield stress sig_1. */
  int32_t curve_type;
  int32_t ncurves;
  double curve_eps[2];                 /* plastic strains of points 1, 2 */
  double curve_temp[W3D_MAX_CURVES];   /* seg_curves_value              */
  double curve_e[W3D_MAX_CURVES];      /* seg_curves_ym                 */
  double curve_nu[W3D_MAX_CURVES];     /* seg_curves_nu                 */
  double curve_alpha[W3D_MAX_CURVES];  /* seg_curves_alpha              */
  double curve_sig[W3D_MAX_CURVES][2]; /* stresses of points 1, 2       */
} w3d_material;

/*
 * The mesh and its element blocking.  Arrays are read during w3d_create and
 * not retained.
 */
typedef struct w3d_model {
  int32_t abi_version;      /* W3D_ABI_VERSION                              */
  int32_t nonode;           /* number of structure nodes                    */
  int32_t noelem;           /* number of elements                           */
  const double* c;          /* (3*nonode) x,y,z of node n at c[3(n-1)+k]    */
  const int32_t* incid;     /* packed element incidences, 1-based nodes     */
  const int32_t* incmap;    /* (noelem) 1-based start of element e in incid */
  const int32_t* etype;     /* (noelem) element type                        */
  const int32_t* int_order; /* (noelem) integration order, iprops(5)        */
  const int32_t* matnum;    /* (noelem) 1-based index into mats             */
  const int32_t* bbar;      /* (noelem) mean-dilatation B-bar (hex8 only)   */
  const int32_t* geonl;     /* (noelem) finite strains and rotations, the
                               reference's lprops(18); NULL: all small     */
  const double* angles;     /* crystal Euler angles: crystal c of element e
                               at angles[3*(angle_stride*(e-1) + c) + k];
                               may be NULL when no element uses mm10        */
  int32_t angle_stride;     /* angle triplets per element (0 means 1)       */
  int32_t nmat;
  const w3d_material* mats;
  int32_t nelblk;           /* >0: user blocking table below; 0: automatic  */
  const int32_t* elblks;    /* (4*nelblk) span, felem, owner, unused        */
  int32_t auto_size;        /* automatic block size (inelbk.f:219)          */
  double eps_bbar;          /* B-bar stabilisation, indypm.f:713 (0.0)      */
  /* Skewed constraints, the reference's trn / trnmat (main_data): node n
     with trn[n-1] != 0 has its displacements, forces and stiffness rows in
     the constraint-compatible system u_c = T u_g, T = trnmat(n)%mat stored
     column-major at trnmat[9(n-1) + r + 3c] (row r, column c).  The
     library rotates element displacements to global coordinates (trnvec,
     dupstr.f:153-160), element forces back (trnvecs, rknifv.f:75-82) and
     element stiffnesses to T Ke T^T (trnmtx, rktstf.f:178-186).  Both NULL:
     no transformations. */
  const int32_t* trn;       /* (nonode) 0/1                                 */
  const double* trnmat;     /* (9*nonode)                                   */
} w3d_model;

typedef struct w3d_ctx w3d_ctx;

/* Thread-local message of the last failing call (never NULL). */
const char* w3d_last_error(void);

/* Library version string. */
const char* w3d_version(void);

/*
 * Build the device-resident model on HIP device `device`: blocking
 * (inelbk.f:282 rules for automatic blocking, or the given table checked
 * against the same compatibility rules), connectivity, dof destinations
 * (edest), inverse incidences (setup.f:160-190), element constants and the
 * zero initial state.
 */
int w3d_create(w3d_ctx** ctx, const w3d_model* model, int device);
int w3d_destroy(w3d_ctx* ctx);

/* Blocking result: nelblk and the (4, nelblk) elblks table. */
int w3d_get_blocking(const w3d_ctx* ctx, int32_t* nelblk, int32_t* elblks);
/* Per-block sizes: span, ngp, hist size, totdof, utsz. */
int w3d_block_info(const w3d_ctx* ctx, int32_t blk, int32_t* info5);

/*
 * Time increment of the current load step (the reference's dt, main_data),
 * which mm10 uses for its rates and its hardening-rate predictor.  The
 * default is 1.0.
 */
int w3d_set_time_increment(w3d_ctx* ctx, double dt);

/*
 * Temperatures of the current load step, the reference's main_data
 * temper_nodes / temper_elems (totals at the start of the step) and
 * dtemp_nodes / dtemp_elems (the step's imposed changes); any of the four
 * may be NULL (zero).  Each element gathers node value + element value
 * (gadtemps, gatemps, gatemps.f:63,117), every Gauss point interpolates them
 * with the shape functions (gauss_pt_temps, gp_temp_eps.f:17), and, when
 * incremental != 0 (the reference's `temperatures` flag: temperature
 * changes imposed in this step), the thermal strain increment
 * alpha(T n+1) (T n+1 - alpha_zero) - alpha(T n) (T n - alpha_zero)
 * (gp_temp_eps, gp_temp_eps.f:163) is taken from the strain increment the
 * material sees (drive_01_update rstgp1.f:804-819, drive_10_update
 * :4400-4447); the stored strains stay those of the displacements.
 * Temperature-dependent mm01 curves read the point temperatures whether or
 * not incremental is set.  Replaces drive_eps_sig_internal_forces.f:
 * 1756-1790 (the gathers of dupstr_blocked).  The values are copied; they
 * hold until the next call.
 */
int w3d_set_temperatures(w3d_ctx* ctx, const double* temper_nodes,
                         const double* dtemp_nodes,
                         const double* temper_elems,
                         const double* dtemp_elems, int32_t incremental);

/*
 * Strain/stress recovery, material update and internal forces for every
 * block, then the deterministic assembly of the element forces into the
 * structure vector ifv (length 3*nonode).  u is the displacement at n, du the
 * increment n -> n+1 (both 3*nonode, structure dof order).
 *   iter = 0 is the reference's pseudo-update (linear-elastic [D] unless
 *   extrapolated != 0, rstgp1.f:835; for mm10 unless extrapolated != 0 and
 *   step > 1, rstgp1.f:4453); history is scattered for iter > 0, and for
 *   mm10 always, since its [D] lives in the history (rplstr.f:64-66).
 * energy / plastic_work receive the block-summed internal energy and plastic
 * work (rstgp1_b); either may be NULL.
 * Returns W3D_CUT_STEP when any Gauss point asks for a step cut.
 */
int w3d_recover(w3d_ctx* ctx, int32_t step, int32_t iter, int32_t extrapolated,
                const double* u, const double* du, double* ifv,
                double* energy, double* plastic_work);
int w3d_recover_dev(w3d_ctx* ctx, int32_t step, int32_t iter,
                    int32_t extrapolated, const double* d_u,
                    const double* d_du, double* d_ifv, double* energy,
                    double* plastic_work);

/*
 * The per-block interface, the counterpart of do_nleps_block /
 * do_nlek_block called inside the reference's !$OMP PARALLEL DO over
 * blocks (drive_eps_sig_internal_forces.f:160-180, tanstf.f:80-87):
 *
 *   w3d_set_displacements(ctx, u, du)           once per recovery pass
 *   w3d_recover_block(ctx, blk, ...)            any thread, any order
 *   w3d_assemble_ifv(ctx, ifv, &sum, &nterm)    after every block (addifv)
 *   w3d_tangent_block(ctx, blk, ...)            any thread, any order
 *
 * w3d_recover_block and w3d_tangent_block may run concurrently from many
 * host threads for distinct blocks of one context: each call works on its
 * own block's state, flags and work items, on a stream of the context's
 * pool, and returns when its block is done.  They must not overlap the
 * context's other calls (whole-pass recovery, commit, rollback,
 * w3d_set_displacements, assembly), as the reference's block loop does not
 * overlap its update or assembly.
 *
 * w3d_set_displacements copies u (may be NULL: zero) and du (3*nonode) to
 * the device after the context's earlier work has ended.
 * w3d_recover_block recovers block blk (1-based) with those displacements
 * and leaves its element forces in the block's einfvec; block_energy and
 * block_plastic_work receive the block's internal energy and plastic work
 * increments summed in element order (rstgp1_b; zero for iter = 0) and
 * cut the block's material_cut_step (also the W3D_CUT_STEP status).  Any
 * output pointer may be NULL.
 * w3d_assemble_ifv gathers every block's element forces into ifv (3*nonode,
 * may be NULL) in a fixed order, and gives addifv's sum of |element force|
 * terms over the blocks' last recoveries (sum_ifv) and their number
 * (num_term_ifv), the input of the convergence test on the largest residual
 * (cvtest.f test 4).
 */
int w3d_set_displacements(w3d_ctx* ctx, const double* u, const double* du);
int w3d_recover_block(w3d_ctx* ctx, int32_t blk, int32_t step, int32_t iter,
                      int32_t extrapolated, double* block_energy,
                      double* block_plastic_work, int32_t* cut);
int w3d_assemble_ifv(w3d_ctx* ctx, double* ifv, double* sum_ifv,
                     int64_t* num_term_ifv);
int w3d_tangent_block(w3d_ctx* ctx, int32_t blk, int32_t step, int32_t iter,
                      int32_t precision);
/* Blocks first_blk .. first_blk+nblk-1 in one call, on one stream: with du
 * non-NULL the displacements are set first (w3d_set_displacements), with
 * du NULL those of the last w3d_set_displacements are used.  Returns
 * W3D_CUT_STEP when any of them asks for a cut. */
int w3d_recover_blocks(w3d_ctx* ctx, int32_t first_blk, int32_t nblk,
                       int32_t step, int32_t iter, int32_t extrapolated,
                       const double* u, const double* du);
/* addifv's sums of the last whole-pass recovery that returned its energy
 * (w3d_recover / w3d_recover_dev / w3d_iteration_dev with an energy or
 * plastic-work pointer): sum of |element force| terms and their number. */
int w3d_get_ifv_sums(const w3d_ctx* ctx, double* sum_ifv,
                     int64_t* num_term_ifv);

/*
 * Block tangent stiffness Ke = sum_gp B^T D B w |J| for every block, from the
 * [D] left by the last recovery (mm01: cep_blocks; mm10: history1(1:36)).
 * precision selects W3D_FP64 or W3D_FP32 (MFMA).
 */
int w3d_tangent(w3d_ctx* ctx, int32_t step, int32_t iter, int32_t precision);
/* blocks first_blk .. first_blk+nblk-1 in one call (the whole model: one
 * launch per kernel group); every block of a pass takes the same
 * precision */
int w3d_tangent_blocks(w3d_ctx* ctx, int32_t first_blk, int32_t nblk,
                       int32_t step, int32_t iter, int32_t precision);

/*
 * Sparsity of the assembled upper triangle for the absolute constraints in
 * cstmap (3*nonode, nonzero = constrained): dof_map (assemble_code.f:561),
 * count_profile_symmetric (:20) and build_col_sparse_symm (:260).
 * Outputs neqns and ncoeff (terms strictly right of the diagonal).
 */
int w3d_pattern(w3d_ctx* ctx, const int32_t* cstmap, int32_t* neqns,
                int64_t* ncoeff);
/* Copy the pattern out: dof_eqn_map (3*nonode), eqn_node_map (neqns),
   k_ptrs (neqns, row counts), k_indexes (ncoeff, 1-based columns).  Any
   pointer may be NULL. */
int w3d_get_pattern(const w3d_ctx* ctx, int32_t* dof_eqn_map,
                    int32_t* eqn_node_map, int32_t* k_ptrs,
                    int32_t* k_indexes);

/*
 * Row-by-row assembly of the current block tangents into k_diag (neqns) and
 * k_coeffs (ncoeff) following assem_by_row (assemble_code.f:659).
 */
int w3d_assemble_vss(w3d_ctx* ctx, double* k_diag, double* k_coeffs);
int w3d_assemble_vss_dev(w3d_ctx* ctx, double* d_k_diag, double* d_k_coeffs);

/* Asymmetric assembly, the solution parameter "assembly asymmetric on"
 * (main_data asymmetric_assembly):
 *  - mm10 keeps the tangent it forms unsymmetrised (mm10_a.f:1165);
 *  - w3d_tangent (FP64 only, whole model) forms the full Ke of every
 *    element, bdbt_asym (gptns1.f:1513): word col*T + row of an element
 *    holds B(row)^T D B(col), D the point's history words 1-36
 *    (drive_10_cnst, gptns1.f:1371) or mm01's cep (extract_D_symmetric);
 *  - w3d_assemble_csr sums them into the values of the full CSR pattern
 *    of w3d_get_csr_pattern (per row: the diagonal and every coupled
 *    equation, ascending 0-based columns), row srow gathering the words
 *    (erow-1)*T + ecol of each element on its node in inverse-incidence
 *    order, assem_by_row_asymmetric / assem_a_row_asym
 *    (assemble_code.f:1050) -- the CSR that convert_vss_csr hands the
 *    asymmetric solver (drive_assemble_solve.f:757);
 *  - the symmetric paths (w3d_assemble_vss*, w3d_iteration_dev, the
 *    batched server) refuse the context while it is on; finite-strain
 *    blocks are refused (W3D_ERR_UNSUPPORTED).
 * Switching clears the last tangent.  w3d_get_block_ek_full copies block
 * blk's full Ke (span * T * T words, element after element). */
int w3d_set_asymmetric(w3d_ctx* ctx, int32_t on);
int w3d_get_csr_pattern(const w3d_ctx* ctx, int64_t* nnz, int64_t* rp,
                        int32_t* col);
int w3d_assemble_csr(w3d_ctx* ctx, double* val);
int w3d_assemble_csr_dev(w3d_ctx* ctx, double* d_val);
int w3d_get_block_ek_full(const w3d_ctx* ctx, int32_t blk, double* out,
                          int64_t n_out);
/* The nonsymmetric solve of the asymmetric CSR of the last
 * w3d_assemble_csr, in place of drive_pardiso_asym (drive_pardiso_asym.f,
 * Pardiso mtype 1, iparm(12) = 0: A x = b with A the rows as assembled):
 * Jacobi-preconditioned BiCGStab from x = 0, FP64, to |r| <= rtol |b|.
 * iters/relres as w3d_solve_pcg; W3D_ERR_NUMERIC after maxit iterations or
 * on a breakdown (a zero (rh, r) or (rh, v)), x holding the last iterate.
 * Its work vectors are allocated with the CSR gather map, once per
 * pattern. */
int w3d_solve_bicgstab(w3d_ctx* ctx, const double* rhs, double* x,
                       double rtol, int32_t maxit, int32_t* iters,
                       double* relres);

/*
 * Jacobi-preconditioned conjugate gradients, FP64, on the device, for the
 * matrix of the last w3d_assemble_vss call (its outputs may be NULL to
 * assemble on the device only): solves K x = rhs over the neqns free
 * equations.  x holds the initial guess on entry and the solution on
 * return.  Stops when ||rhs - K x|| <= rtol ||rhs|| (W3D_OK) or after maxit
 * iterations (W3D_ERR_NUMERIC, x the last iterate); iters and relres (either
 * may be NULL) receive the iterations and the final relative residual.
 * Replaces the reference's assembled-equation solve (drive_assemble_solve.f
 * :654 -> drive_pardiso.f:11; eqn_solve in mnralg.f:503).
 */
int w3d_solve_pcg(w3d_ctx* ctx, const double* rhs, double* x, double rtol,
                  int32_t maxit, int32_t* iters, double* relres);
/* Device-pointer form: the matrix is the d_k_diag / d_k_coeffs pair of a
   w3d_assemble_vss_dev call. */
int w3d_solve_pcg_dev(w3d_ctx* ctx, const double* d_k_diag,
                      const double* d_k_coeffs, const double* d_rhs,
                      double* d_x, double rtol, int32_t maxit,
                      int32_t* iters, double* relres);

/*
 * w3d_solve_pcg_dev for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * It has no counterpart in the reference.  Same arguments and results as
 * w3d_solve_pcg_dev, but its work vectors live in a static device array of
 * the library (4 * 2^19 + 1032 doubles, about 16.8 MB per device, held
 * until the process ends) instead of the context's pattern-time workspace;
 * it calls no allocator.  Limit: at most 2^19 = 524,288 equations
 * (W3D_ERR_ARG beyond).  Every call on one device shares that array:
 * calls that overlap on one device (from several threads or contexts)
 * race and must be serialised by the caller.
 */
int w3d_solve_pcg_static(w3d_ctx* ctx, const double* d_k_diag,
                         const double* d_k_coeffs, const double* d_rhs,
                         double* d_x, double rtol, int32_t maxit,
                         int32_t* iters, double* relres);

/* n+1 -> n roll-over of stresses, strains and history (update.f:15). */
int w3d_commit_step(w3d_ctx* ctx);
/* Discard the n+1 state (adaptive step restart): n+1 <- n. */
int w3d_rollback(w3d_ctx* ctx);

/*
 * Host copy of one block's state in the reference's layout (see the
 * W3D_* selectors).  n_out is the capacity of out in doubles; W3D_ESTIFF of
 * an FP32 tangent is widened to double.
 */
int w3d_get_block_state(const w3d_ctx* ctx, int32_t blk, int32_t which,
                        double* out, int64_t n_out);
/* Overwrite one block's n-state (restart / initial state).  which is
   W3D_STRESS_N, W3D_STRAIN_N or W3D_HIST_N. */
int w3d_set_block_state(w3d_ctx* ctx, int32_t blk, int32_t which,
                        const double* in, int64_t n_in);

/* Wait for all device work of the context. */
int w3d_synchronize(w3d_ctx* ctx);

/*
 * One Newton iteration of the element-block path on device buffers, as one
 * HIP graph launch: w3d_recover_dev (all blocks, into d_ifv), w3d_tangent
 * (all blocks) and w3d_assemble_vss_dev into (d_k_diag, d_k_coeffs), with
 * the same results.  Replaces the call sequence of a Newton iteration of
 * mnralg.f (drive_eps_sig_internal_forces :358/:1102, tanstf :367/:439,
 * assem_by_row via eqn_solve :503).  The graph is captured on first use for
 * each (step 1 or later, iteration 0 or later, extrapolated, precision,
 * energy wanted, caller buffers, time increment) and replayed afterwards;
 * up to 8 are kept.  Needs w3d_pattern first.  Returns like w3d_recover.
 */
int w3d_iteration_dev(w3d_ctx* ctx, int32_t step, int32_t iter,
                      int32_t extrapolated, const double* d_u,
                      const double* d_du, double* d_ifv, int32_t precision,
                      double* d_k_diag, double* d_k_coeffs, double* energy,
                      double* plastic_work);

/*
 * Autotuner (no reference counterpart: the reference's block size is the
 * fixed mxvl): times the block tangent of every kernel group with work
 * items of 8, 16, 32 and 64 elements and keeps the fastest, and the
 * persistent mm10 kernel with 2, 4 and 8 workgroups per compute unit.  Run
 * it after a recovery of a representative iteration (mm10 is timed on the
 * increments in place).  choices receives (group key, elements) pairs, then
 * (-1, mm10 workgroups) when the model has mm10 blocks; *n_choices their
 * number of words.  Captured iterations are dropped.
 */
int w3d_autotune(w3d_ctx* ctx, int32_t precision, int32_t* choices,
                 int32_t capacity, int32_t* n_choices);

/*
 * Per-phase device timing, the counterpart of the reference's thyme()
 * accumulators (thyme.f:15): when enabled, every kernel launch of a phase is
 * bracketed by HIP events on the context's stream.  Phases: 0 recovery
 * (kinematics, mm01 and element forces), 1 element-force gather, 2 energy
 * reduction, 3 tangent, 4 assembly, 5 the mm10 material update.
 */
enum {
  W3D_PHASE_RECOVER = 0,
  W3D_PHASE_IFV = 1,
  W3D_PHASE_ENERGY = 2,
  W3D_PHASE_TANGENT = 3,
  W3D_PHASE_ASSEMBLE = 4,
  W3D_PHASE_MM10 = 5,
  W3D_NUM_PHASES = 6
};
int w3d_set_profiling(w3d_ctx* ctx, int32_t on);
/* launches and total device milliseconds of one phase since the last reset */
int w3d_get_phase_time(const w3d_ctx* ctx, int32_t phase, int64_t* launches,
                       double* total_ms);
int w3d_reset_phase_times(w3d_ctx* ctx);

/*
 * Batched serving of independent problems (no reference counterpart; the
 * reference solves one model per process).  A problem is a quasi-static
 * solve on one context: nsteps load steps, each imposing the increments
 * du0 gives on the constrained dofs of cstmap (nonzero = constrained), with
 * Newton iterations of recovery, block tangent, assembly and a Jacobi-PCG
 * solve of the free equations until the residual (-ifv on the free dofs)
 * satisfies |R| <= tol * max(|F_c|, tol_abs), F_c the internal forces on
 * the constrained dofs.  Step 1 starts from du0 (its free entries are the
 * initial guess), every later step from the previous step's converged
 * increment (extrapolation).  With adaptive on, an increment whose
 * recovery asks for a cut (or that reaches max_iters) is restored and
 * split into four sub-increments, recursively down to 1/16 of the step,
 * with the stack, the first-sub-increment restart and the extrapolation
 * scales of the reference's adaptive solution (mnralg.f, adapt_check.f),
 * from step 1 on.  Mesh, state, displacements and solver vectors stay on
 * the device.
 *
 * w3d_batch_advance runs one Newton iteration of every running problem of
 * a batch: each kernel of the iteration is launched once for the whole
 * batch where its work spans problems (the mm10 update, the residual, the
 * Jacobi-PCG iterations, the displacement update) and per problem for the
 * rest, on the batch's stream.  A problem's arithmetic does not depend on
 * the other problems of the batch: it gives bitwise the same results alone
 * and in any mix.  Problems may join or leave between calls.  The contexts
 * of the problems must not be used by other calls during an advance.
 */
typedef struct w3d_problem w3d_problem;
typedef struct w3d_batch w3d_batch;

typedef struct w3d_newton_opts {
  int32_t nsteps;     /* load steps                                     */
  int32_t max_iters;  /* Newton iterations per (sub-)increment          */
  int32_t precision;  /* block tangent: W3D_FP64 or W3D_FP32            */
  int32_t pcg_maxit;  /* Jacobi-PCG iterations per solve                */
  int32_t adaptive;   /* 1: split failing increments (adapt_check.f)    */
  int32_t reserved;   /* 0                                              */
  double tol;         /* relative residual of convergence               */
  double tol_abs;     /* floor of the reaction norm in the test         */
  double pcg_rtol;    /* |r| <= pcg_rtol |b| ends a linear solve        */
} w3d_newton_opts;

enum {
  W3D_PROBLEM_RUNNING = 0,
  W3D_PROBLEM_STEP_DONE = 1, /* a step converged in the last advance    */
  W3D_PROBLEM_DONE = 2,      /* every step converged                    */
  W3D_PROBLEM_FAILED = 3     /* iteration limit or a material failure   */
};

/* Sets the context's pattern from cstmap and allocates the problem's
 * device vectors; du0 has 3*nonode words. */
int w3d_problem_create(w3d_ctx* ctx, const int32_t* cstmap, const double* du0,
                       const w3d_newton_opts* opts, w3d_problem** out);
int w3d_problem_destroy(w3d_problem* p);
/* status (W3D_PROBLEM_*), current step (1-based) and Newton iteration */
int w3d_problem_status(const w3d_problem* p, int32_t* status, int32_t* step,
                       int32_t* iter);
/* Record of converged step (1-based): Newton iterations (all attempts),
 * Jacobi-PCG iterations summed over them, adaptive cuts, the sum of the
 * internal forces on the loaded dofs (constrained dofs with a nonzero
 * increment), and the last residual and reaction norms. */
int w3d_problem_step_record(const w3d_problem* p, int32_t step,
                            int32_t* newton_iters, int32_t* pcg_iters,
                            int32_t* cuts, double* load, double* residual,
                            double* reaction);
/* displacements at the end of the last converged step (3*nonode) */
int w3d_problem_get_u(const w3d_problem* p, double* u);
/* device bytes the problem's context and vectors hold */
int w3d_problem_device_bytes(const w3d_problem* p, int64_t* bytes);

int w3d_batch_create(int32_t device, w3d_batch** out);
int w3d_batch_destroy(w3d_batch* b);
/* One Newton iteration of each of the n problems that are still running;
 * status[k] receives problem k's status after it. */
int w3d_batch_advance(w3d_batch* b, w3d_problem* const* problems, int32_t n,
                      int32_t* status);
/* Device milliseconds of the batch's advances so far, per phase: ms[0]
 * recovery (all passes), ms[1] the batched mm10 inside it, ms[2] residual,
 * ms[3] tangents and assemblies, ms[4] the batched Jacobi-PCG (with its
 * host checks), ms[5] the update; the number of advances and of batched
 * PCG iterations. */
int w3d_batch_phase_times(const w3d_batch* b, double* ms, int64_t* advances,
                          int64_t* pcg_iterations);

#ifdef __cplusplus
}
#endif

#endif /* W3D_H */
