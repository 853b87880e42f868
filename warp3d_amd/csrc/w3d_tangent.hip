// w3d_tangent.hip -- block tangent Ke = sum_gp B^T (D w |J|) B on the
// matrix cores (gfx950 MFMA), FP32 (v_mfma_f32_16x16x4f32, FP32
// accumulation) or FP64 (v_mfma_f64_16x16x4f64).
//
// Reference: tanstf.f:119 do_nlek_block, rktstf.f:15, gptns1.f:16 gptns1,
// gptns1.f:1174 extract_D_symmetric, gptns1.f:1628 bdbtgen, gpmas1.f:399
// adjust_cnst.  bdbtgen forms, per point, BD = D B and then the upper
// triangle Ke(row,col) += sum_r B(r,col) BD(r,row).  Here the sum over
// points and strain components is one contraction over k = (r, point):
//   Ke = A B   with  A(row, k) = BD(r, row) at point p,  B(k, col) = B(r, col)
// taken 4 k at a time by the 16x16x4 MFMA on 16x16 tiles of the upper
// triangle.  A 4-k step is one strain component r at four points, so each
// lane (point q = lane/16 of the step, dof column lane%16 of every tile)
// builds its own A and B operands in registers: the three gradient
// components of its node at its point, the B column (with the mean
// dilatation terms for B-bar) and the six D B values, reused over the six
// strain components.  Columns are the element dofs renumbered by node rank
// (3*rank + component), so the accumulators store directly into the ranked
// row-major element layout (ek_word) in runs of 16 words.
//
// A wave takes a batch of elements (tan_batch(G) of them, lanes = element x
// point for the set-up): it gathers their nodes and coordinates, forms the
// Jacobians, inverses and B-bar terms of all their points at once and
// loads their [D] words coalesced (one point per lane), then contracts the
// elements one after the other on the matrix cores.
//
// Finite strains (GEONL; gptns1.f:307 gptns1_a, :1735 ctran1, kg1.f:16
// kgstiff): the configuration is X + u + du of the last recovery; [D] is
// rotated to the Cauchy frame with the stored [R n+1] and corrected by the
// initial-stress matrix (FP64, per point); the geometric stiffness
// gtg(a, b) = sum_p (g_a^T S g_b) w |J| of every node pair is a second,
// smaller contraction on the matrix cores (k = (point, component)) whose
// NN x NN result is added to the three component blocks as Ke is stored.
#include <hip/hip_runtime.h>

#include "w3d_geonl.h"
#include "w3d_internal.h"

namespace w3d {

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row of output value v held by lane l (column is l & 15)
  __device__ static int row(int l, int v) { return 4 * (l >> 4) + v; }
};
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row(int l, int v) { return (l >> 4) + 4 * v; }
};

// Wave-level synchronisation: the lanes of one wave exchange data through
// their own LDS slab, so the wave only has to order its own LDS traffic.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The B columns of a node's three components (x, y, z) from its gradient
// (bx, by, bz) and, for B-bar, its volume-averaged gradient vol[3 j + c]:
// bc[c][r] is row r (xx, yy, zz, xy, yz, xz) of component c's column.
template <bool BBAR, typename T>
__device__ inline void column_xyz(T bx, T by, T bz, const T* vol, int j,
                                  T alpha, T beta_b, T (&bc)[3][6]) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 6; ++r) bc[c][r] = 0;
  bc[0][0] = bx; bc[0][3] = by; bc[0][5] = bz;
  bc[1][1] = by; bc[1][3] = bx; bc[1][4] = bz;
  bc[2][2] = bz; bc[2][4] = by; bc[2][5] = bx;
  if (BBAR) {
    const T vx = vol[3 * j], vy = vol[3 * j + 1], vz = vol[3 * j + 2];
    bc[0][0] = alpha * bx + beta_b * vx;
    bc[0][1] = bc[0][2] = beta_b * (vx - bx);
    bc[1][0] = bc[1][2] = beta_b * (vy - by);
    bc[1][1] = alpha * by + beta_b * vy;
    bc[2][0] = bc[2][1] = beta_b * (vz - bz);
    bc[2][2] = alpha * bz + beta_b * vz;
  }
}

// Waves per SIMD the register allocation aims at: three for the 8-node
// brick's small-strain tangent (168 registers, no spills; at four its
// 128-register cap spilled 14 words of the batch set-up and the FP64
// tangent ran 11 % slower, the FP32 one the same); two for the others,
// whose LDS holds the CU to fewer.
__host__ __device__ constexpr int tan_waves_per_eu(int nnode, bool geonl) {
  return nnode == 8 && !geonl ? 3 : 2;
}

template <int NN, int G, bool BBAR, typename T, bool GEONL = false,
          bool TRN = false>
__global__ void __launch_bounds__(64 * tan_mfma_waves(NN))
    __attribute__((amdgpu_waves_per_eu(tan_waves_per_eu(NN, GEONL))))
    k_tangent_mfma(DevModel M, const WorkItem* __restrict__ work,
                   T* __restrict__ ek_out) {
  constexpr int W = tan_mfma_waves(NN);   // waves per workgroup
  constexpr int EB = tan_batch(G, (int)sizeof(T));  // elements of a batch
  constexpr int TD = 3 * NN;              // element dofs
  constexpr int NTL = (TD + 15) / 16;     // 16-wide tiles per side
  constexpr int NUP = NTL * (NTL + 1) / 2;  // upper tiles
  constexpr int NPG = (G + 3) / 4;        // 4-point groups (one k step each)
  using MF = Mfma<T>;
  using v4 = typename MF::v4;

  // shape-function derivatives: FP64 for the Jacobians, and in the MFMA
  // type for the gradients (the same array when that type is FP64)
  constexpr bool F64 = sizeof(T) == 8;
  __shared__ double s_dn64[3][G][NN];
  __shared__ T s_dnt[F64 ? 1 : 3][G][NN];
  T(*s_dn)[G][NN] = F64 ? reinterpret_cast<T(*)[G][NN]>(s_dn64) : s_dnt;
  __shared__ double s_w[G];
  __shared__ double s_detw[W][EB][G];
  __shared__ T s_gam[W][EB][G][9];
  __shared__ T s_vol[W][BBAR ? EB : 1][3 * NN];
  // the batch's scaled [D] words (S3 on) share each wave's slab with the
  // nodal coordinates (S1-S2 only): one workgroup more per CU.  The slabs
  // are per wave, as the waves run their stages independently.
  constexpr size_t kDBytes = sizeof(T) * EB * G * 21;
  constexpr size_t kCeBytes = sizeof(double) * EB * 3 * NN;
  constexpr size_t kSlab = (kDBytes > kCeBytes ? kDBytes : kCeBytes) / 8;
  __shared__ double s_dce[W][kSlab];
  __shared__ int s_node[W][EB][NN];
  __shared__ int s_inv[W][EB][NN];
  // finite strains: the Cauchy stresses of the batch's points, and the
  // geometric stiffness of the current element's node pairs
  constexpr int GW = GEONL ? W : 1;
  __shared__ T s_cs[GW][GEONL ? EB : 1][GEONL ? G : 1][6];
  __shared__ double s_gx[GW][GEONL ? NN * NN : 1];

  const WorkItem wi = work[blockIdx.x];
  const BlockDesc bd = M.blocks[wi.blk];
  const GaussTable* gt = M.tables + bd.table;
  const int tid = threadIdx.x;
  const int wv = tid >> 6;   // wave of this lane
  // this wave's slab as [D] words and as coordinates
  T(*s_dw)[G][21] = reinterpret_cast<T(*)[G][21]>(&s_dce[wv][0]);
  double(*s_cew)[3][NN] = reinterpret_cast<double(*)[3][NN]>(&s_dce[wv][0]);
  const int l = tid & 63;    // lane in the wave
  const int q = l >> 4;      // point of the lane inside a 4-point group
  const int cres = l & 15;   // column residue inside a tile
  const bool tet = bd.etype == W3D_TET10;
  const T alpha = (T)((2.0 + M.eps_bbar) * (1.0 / 3.0));
  const T beta_b = (T)((1.0 - M.eps_bbar) * (1.0 / 3.0));
  const int64_t SG = (int64_t)bd.span * G;
  const double* cep_b = M.cep + 21 * bd.q_off;
  // the point of this lane in the batch set-up: element pe, point pg
  const int pe = l / G, pgp = l % G;

  // the lane's component in each tile's dof column
  int cmp[NTL];
#pragma unroll
  for (int t = 0; t < NTL; ++t) cmp[t] = (16 * t + cres) % 3;
  // small elements (at most 3 upper tiles): the word of each accumulator
  // value in the ranked upper triangle, or -1 below the diagonal or past
  // the last dof
  constexpr bool kPreOff = NUP <= 3;
  int soff[kPreOff ? 4 * NUP : 1];
  if (kPreOff) {
#pragma unroll
    for (int u = 0; u < NUP; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int rw = 16 * tile_i(u) + MF::row(l, v);
        const int cl = 16 * tile_j(u) + cres;
        soff[kPreOff ? 4 * u + v : 0] =
            (rw <= cl && cl < TD) ? ek_word(TD, rw, cl) : -1;
      }
  }

  for (int k = tid; k < 3 * G * NN; k += 64 * W) {
    const int d = k / (G * NN), rem = k % (G * NN);
    const double v = gt->dn[d][rem / NN][rem % NN];
    s_dn64[d][rem / NN][rem % NN] = v;
    if (!F64) s_dnt[d][rem / NN][rem % NN] = (T)v;
  }
  if (tid < G) s_w[tid] = gt->w[tid];
  __syncthreads();

  // each wave takes batches of EB consecutive elements of the work item
  for (int b0 = wv * EB; b0 < wi.count; b0 += W * EB) {
    const int nb = wi.count - b0 < EB ? wi.count - b0 : EB;
    const int ie0 = wi.first + b0;
    wave_sync();
    // S1: nodes and coordinates of the batch (lanes = element x node)
    for (int k = l; k < nb * NN; k += 64) {
      const int e = k / NN, j = k % NN;
      const int node = M.conn[bd.conn_off + (int64_t)(ie0 + e) * NN + j];
      s_node[wv][e][j] = node;
      double x[3] = {M.coords[3 * node], M.coords[3 * node + 1],
                     M.coords[3 * node + 2]};
      if (GEONL) {  // dptstf: X + u + du (skewed nodes' rotated to global)
        double uu[3], dd[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          uu[c] = M.u_t[3 * node + c];
          dd[c] = M.du_t[3 * node + c];
        }
        const int ti = TRN ? M.trn_idx[node] : -1;
        if (ti >= 0) {
          trn_apply_t(M.trn_mat + 9 * ti, uu);
          trn_apply_t(M.trn_mat + 9 * ti, dd);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c] = x[c] + uu[c] + dd[c];
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) s_cew[e][c][j] = x[c];
    }
    wave_sync();
    const bool pact = l < nb * G;   // the lane has a point of the batch
    // S2: node ranks; the Jacobian at every point (jacob1) and its inverse
    for (int k = l; k < nb * NN; k += 64) {
      const int e = k / NN, j = k % NN;
      const int mine = s_node[wv][e][j];
      int r = 0;
#pragma unroll
      for (int bb = 0; bb < NN; ++bb) {
        const int o = s_node[wv][e][bb];
        r += (o < mine || (o == mine && bb < j)) ? 1 : 0;
      }
      s_inv[wv][e][r] = j;
    }
    if (pact) {
      const int e = pe, g = pgp;
      double j11 = 0, j12 = 0, j13 = 0, j21 = 0, j22 = 0, j23 = 0, j31 = 0,
             j32 = 0, j33 = 0;
      for (int j = 0; j < NN; ++j) {
        const double x = s_cew[e][0][j], y = s_cew[e][1][j],
                     z = s_cew[e][2][j];
        const double a = s_dn64[0][g][j], b = s_dn64[1][g][j],
                     c = s_dn64[2][g][j];
        j11 = j11 + a * x; j12 = j12 + a * y; j13 = j13 + a * z;
        j21 = j21 + b * x; j22 = j22 + b * y; j23 = j23 + b * z;
        j31 = j31 + c * x; j32 = j32 + c * y; j33 = j33 + c * z;
      }
      const double c1 = j22 * j33 - j23 * j32;
      const double c2 = j21 * j33 - j23 * j31;
      const double c3 = j21 * j32 - j22 * j31;
      const double dj = j11 * c1 - j12 * c2 + j13 * c3;
      // one reciprocal for the nine cofactors (jacob1 divides each; the
      // difference is one rounding of the inverse, far below the FP64
      // tangent's tolerance)
      const double rdj = 1.0 / dj;
      T* ga = s_gam[wv][e][g];
      ga[0] = (T)(c1 * rdj);
      ga[3] = (T)(-c2 * rdj);
      ga[6] = (T)(c3 * rdj);
      ga[1] = (T)((j32 * j13 - j12 * j33) * rdj);
      ga[4] = (T)((j11 * j33 - j31 * j13) * rdj);
      ga[7] = (T)((j12 * j31 - j11 * j32) * rdj);
      ga[2] = (T)((j12 * j23 - j13 * j22) * rdj);
      ga[5] = (T)((j13 * j21 - j11 * j23) * rdj);
      ga[8] = (T)((j11 * j22 - j12 * j21) * rdj);
      s_detw[wv][e][g] = dj;
    }
    wave_sync();
    // this lane's point: the 21 words of its stored [D], coalesced over
    // the batch (word k of point (i, g) at k S G + i G + g); issued after
    // the Jacobians, which need many registers, and in flight during the
    // B-bar sums; lanes past the batch read its last point (in bounds) and drop it
    const int64_t pq = (int64_t)ie0 * G + (pact ? l : nb * G - 1);
    double pd[21];
#pragma unroll
    for (int k = 0; k < 21; ++k) pd[k] = cep_b[k * SG + pq];
    // S3: B-bar volume terms (lanes = element x node); [D] scaled by
    // w |J| (extract_D_symmetric) and by 1/6 for tetrahedra (adjust_cnst)
    if (BBAR)
      for (int k = l; k < nb * NN; k += 64) {
        const int e = k / NN, j = k % NN;
        T vx = 0, vy = 0, vz = 0;
        double volume = 0;
        for (int g = 0; g < G; ++g) {
          const T* a = s_gam[wv][e][g];
          const T d = (T)s_detw[wv][e][g];
          const T nx = s_dn[0][g][j], ny = s_dn[1][g][j], nz = s_dn[2][g][j];
          vx = vx + (a[0] * nx + a[1] * ny + a[2] * nz) * d;
          vy = vy + (a[3] * nx + a[4] * ny + a[5] * nz) * d;
          vz = vz + (a[6] * nx + a[7] * ny + a[8] * nz) * d;
          volume = volume + s_detw[wv][e][g];
        }
        const T voli = (T)(1.0 / volume);
        s_vol[wv][e][3 * j] = vx * voli;
        s_vol[wv][e][3 * j + 1] = vy * voli;
        s_vol[wv][e][3 * j + 2] = vz * voli;
      }
    if (pact) {
      const int e = pe, g = pgp;
      const double wdj = s_w[g] * s_detw[wv][e][g];
      if (GEONL) {
        // ctran1 (+ Q-bar) with [qn1] of the stored [R n+1] and the
        // Cauchy stress {T} = [qn1] {t n+1}; then adjust_cnst
        const int64_t qp = (int64_t)(ie0 + e) * G + g;
        const double* rr = M.rot_n1 + 9 * bd.q_off + qp;
        const double* sn = M.stress_n1 + 9 * bd.q_off + qp;
        double r9[9], sg[6], q2[36], cs[6], dm[36];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
          for (int c = 0; c < 3; ++c) r9[3 * a + c] = rr[(a + 3 * c) * SG];
#pragma unroll
        for (int k = 0; k < 6; ++k) sg[k] = sn[k * SG];
        geonl_q(r9, 2, q2);
        geonl_qmul(q2, sg, cs);
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
          for (int c = 0; c < 6; ++c) dm[6 * r + c] = wdj * pd[tri(r, c)];
        geonl_ctran(dm, q2, cs, wdj);
#pragma unroll
        for (int r = 0; r < 6; ++r)
#pragma unroll
          for (int c = 0; c <= r; ++c) {
            double v = dm[6 * r + c];
            if (tet) v = (1.0 / 6.0) * v;
            s_dw[e][g][tri(r, c)] = (T)v;
          }
#pragma unroll
        for (int k = 0; k < 6; ++k) s_cs[wv][e][g][k] = (T)cs[k];
      } else {
#pragma unroll
        for (int k = 0; k < 21; ++k) {
          double v = wdj * pd[k];
          if (tet) v = (1.0 / 6.0) * v;
          s_dw[e][g][k] = (T)v;
        }
      }
    }
    wave_sync();
    // S4: the elements of the batch one by one on the matrix cores
    for (int e = 0; e < nb; ++e) {
      const int ie = ie0 + e;
      // the lane's dof column of every tile: node (local index); its
      // component is the per-lane constant cmp[t]
      int jn[NTL];
#pragma unroll
      for (int t = 0; t < NTL; ++t) {
        const int dp = 16 * t + cres;
        jn[t] = dp < TD ? s_inv[wv][e][dp / 3] : -1;
      }
      // the contraction, four points per k step
      v4 acc[NUP];
#pragma unroll
      for (int u = 0; u < NUP; ++u) acc[u] = v4{0, 0, 0, 0};
#pragma unroll 1
      for (int pg = 0; pg < NPG; ++pg) {
        const int p = 4 * pg + q;
        const int pp = p < G ? p : G - 1;
        const T* ga = s_gam[wv][e][pp];
        const T g0 = ga[0], g1 = ga[1], g2 = ga[2], g3 = ga[3], g4 = ga[4],
                g5 = ga[5], g6 = ga[6], g7 = ga[7], g8 = ga[8];
        T col[NTL][6];
#pragma unroll
        for (int t = 0; t < NTL; ++t) {
          const int j = jn[t] < 0 ? 0 : jn[t];
          const T nx = s_dn[0][pp][j], ny = s_dn[1][pp][j],
                  nz = s_dn[2][pp][j];
          const T bx = g0 * nx + g1 * ny + g2 * nz;
          const T by = g3 * nx + g4 * ny + g5 * nz;
          const T bz = g6 * nx + g7 * ny + g8 * nz;
          T c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0;
          if (TRN) {
            // skewed constraints: the column of B T^T, sum_c' T(c, c')
            // B(:, c') over the node's three component columns (trnmtx's
            // T Ke T^T formed inside the contraction)
            T bc[3][6];
            column_xyz<BBAR, T>(bx, by, bz, s_vol[wv][BBAR ? e : 0], j,
                                alpha, beta_b, bc);
            // the component's row of the node's matrix (identity row for
            // a node without one)
            const int ti = M.trn_idx[s_node[wv][e][j]];
            const double* tm = M.trn_mat + 9 * (ti >= 0 ? ti : 0);
            const int ct = cmp[t];
            const T r0 = ti >= 0 ? (T)tm[ct] : (T)(ct == 0 ? 1 : 0);
            const T r1 = ti >= 0 ? (T)tm[ct + 3] : (T)(ct == 1 ? 1 : 0);
            const T r2 = ti >= 0 ? (T)tm[ct + 6] : (T)(ct == 2 ? 1 : 0);
            c0 = r0 * bc[0][0] + r1 * bc[1][0] + r2 * bc[2][0];
            c1 = r0 * bc[0][1] + r1 * bc[1][1] + r2 * bc[2][1];
            c2 = r0 * bc[0][2] + r1 * bc[1][2] + r2 * bc[2][2];
            c3 = r0 * bc[0][3] + r1 * bc[1][3] + r2 * bc[2][3];
            c4 = r0 * bc[0][4] + r1 * bc[1][4] + r2 * bc[2][4];
            c5 = r0 * bc[0][5] + r1 * bc[1][5] + r2 * bc[2][5];
          } else {
            // the column of component cmp[t] (x: bx 0 0 by 0 bz, y: 0 by
            // 0 bx bz 0, z: 0 0 bz 0 by bx), formed by selects on the
            // lane's fixed component rather than by diverging branches
            const bool x = cmp[t] == 0, y = cmp[t] == 1, z = cmp[t] == 2;
            c3 = x ? by : (y ? bx : (T)0);
            c4 = y ? bz : (z ? by : (T)0);
            c5 = x ? bz : (z ? bx : (T)0);
            const T bo = x ? bx : (y ? by : bz);
            if (BBAR) {
              // the volumetric rows: the mean dilatation term of the own
              // gradient on the diagonal row, beta (v - b) on the other two
              const T vo = s_vol[wv][e][3 * j + cmp[t]];
              const T dg = alpha * bo + beta_b * vo;
              const T of = beta_b * (vo - bo);
              c0 = x ? dg : of;
              c1 = y ? dg : of;
              c2 = z ? dg : of;
            } else {
              c0 = x ? bo : (T)0;
              c1 = y ? bo : (T)0;
              c2 = z ? bo : (T)0;
            }
          }
          // points past the rule (the last 4-point group of a rule whose
          // size is not a multiple of 4) contribute nothing
          const bool live = (G % 4 == 0) || p < G;
          col[t][0] = live ? c0 : (T)0;
          col[t][1] = live ? c1 : (T)0;
          col[t][2] = live ? c2 : (T)0;
          col[t][3] = live ? c3 : (T)0;
          col[t][4] = live ? c4 : (T)0;
          col[t][5] = live ? c5 : (T)0;
        }
        const T* dm = s_dw[e][pp];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          // BD(r, dof) = sum_m D(m, r) B(m, dof), bdbtgen order
          const T d0 = dm[tri(0, r)], d1 = dm[tri(1, r)], d2 = dm[tri(2, r)],
                  d3 = dm[tri(3, r)], d4 = dm[tri(4, r)], d5 = dm[tri(5, r)];
          T bdv[NTL];
#pragma unroll
          for (int t = 0; t < NTL; ++t)
            bdv[t] = d0 * col[t][0] + d1 * col[t][1] + d2 * col[t][2] +
                     d3 * col[t][3] + d4 * col[t][4] + d5 * col[t][5];
#pragma unroll
          for (int u = 0; u < NUP; ++u)
            acc[u] = MF::mma(bdv[tile_i(u)], col[tile_j(u)][r], acc[u]);
        }
      }
      if (GEONL) {
        // kgstiff on the matrix cores: one 4-k step per point, k = the
        // three gradient components (and a zero row); lane (component
        // q, node 16 t + cres)
        constexpr int NTG = (NN + 15) / 16;
        constexpr int NUG = NTG * (NTG + 1) / 2;
        v4 accg[NUG];
#pragma unroll
        for (int u = 0; u < NUG; ++u) accg[u] = v4{0, 0, 0, 0};
#pragma unroll 1
        for (int p = 0; p < G; ++p) {
          const T* ga = s_gam[wv][e][p];
          const T* sp = s_cs[wv][e][p];
          const T wdj = (T)(s_w[p] * s_detw[wv][e][p]);
          T av[NTG], bv[NTG];
#pragma unroll
          for (int t = 0; t < NTG; ++t) {
            const int a = 16 * t + cres;
            T gx = 0, gy = 0, gz = 0;
            if (a < NN) {
              if (BBAR) {
                gx = s_vol[wv][e][3 * a];
                gy = s_vol[wv][e][3 * a + 1];
                gz = s_vol[wv][e][3 * a + 2];
              } else {
                const T nx = s_dn[0][p][a], ny = s_dn[1][p][a],
                        nz = s_dn[2][p][a];
                gx = ga[0] * nx + ga[1] * ny + ga[2] * nz;
                gy = ga[3] * nx + ga[4] * ny + ga[5] * nz;
                gz = ga[6] * nx + ga[7] * ny + ga[8] * nz;
              }
            }
            // (S g)_q with S = (xx, yy, zz, xy, yz, xz)
            const T s0 = q == 0 ? sp[0] : q == 1 ? sp[3] : sp[5];
            const T s1 = q == 0 ? sp[3] : q == 1 ? sp[1] : sp[4];
            const T s2 = q == 0 ? sp[5] : q == 1 ? sp[4] : sp[2];
            const T gq = q == 0 ? gx : q == 1 ? gy : gz;
            av[t] = q < 3 ? (s0 * gx + s1 * gy + s2 * gz) * wdj : (T)0;
            bv[t] = q < 3 ? gq : (T)0;
          }
#pragma unroll
          for (int u = 0; u < NUG; ++u)
            accg[u] = MF::mma(av[tile_i(u)], bv[tile_j(u)], accg[u]);
        }
        // the symmetric NN x NN result into the wave's FP64 slab
        wave_sync();
#pragma unroll
        for (int u = 0; u < NUG; ++u) {
          const int I = tile_i(u), J = tile_j(u);
          const int bcol = 16 * J + cres;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int a = 16 * I + MF::row(l, v);
            if (a < NN && bcol < NN) {
              s_gx[wv][a * NN + bcol] = (double)accg[u][v];
              s_gx[wv][bcol * NN + a] = (double)accg[u][v];
            }
          }
        }
        wave_sync();
      }
      // the upper triangle straight from the accumulators, ranked
      // row-major: each store writes four runs of 16 consecutive words
      T* out = ek_out + bd.ek_off + (int64_t)ie * (TD * (TD + 1) / 2);
      if (kPreOff && !GEONL) {
#pragma unroll
        for (int u = 0; u < NUP; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (soff[4 * u + v] >= 0) out[soff[4 * u + v]] = acc[u][v];
        continue;
      }
      // larger elements: the store offsets depend on the lane only and
      // are formed here, per element, rather than held in registers
      // across the element loop
      int ls = l;
      asm volatile("" : "+v"(ls));
#pragma unroll
      for (int u = 0; u < NUP; ++u) {
        const int I = tile_i(u), J = tile_j(u);
        const int cl = 16 * J + (ls & 15);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int rw = 16 * I + MF::row(ls, v);
          if (rw <= cl && cl < TD) {
            T val = acc[u][v];
            if (GEONL) {
              const int ja = s_inv[wv][e][rw / 3], jb = s_inv[wv][e][cl / 3];
              const double gab = s_gx[wv][ja * NN + jb];
              if (TRN) {
                // the geometric stiffness g_ab I of the node pair rotated
                // to T_a (g_ab I) T_b^T
                const int ta = M.trn_idx[s_node[wv][e][ja]];
                const int tb = M.trn_idx[s_node[wv][e][jb]];
                const int ca = rw % 3, cb = cl % 3;
                double dot = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                  const double xa = ta >= 0 ? M.trn_mat[9 * ta + ca + 3 * k]
                                            : (k == ca ? 1.0 : 0.0);
                  const double xb = tb >= 0 ? M.trn_mat[9 * tb + cb + 3 * k]
                                            : (k == cb ? 1.0 : 0.0);
                  dot = dot + xa * xb;
                }
                val = val + (T)(gab * dot);
              } else if (rw % 3 == cl % 3) {
                val = val + (T)gab;
              }
            }
            out[ek_word(TD, rw, cl)] = val;
          }
        }
      }
    }
  }
}

#define W3D_TAN_GROUPS(X) \
  X(8, 8, true)           \
  X(8, 8, false)          \
  X(20, 27, false)        \
  X(20, 8, false)         \
  X(20, 14, false)        \
  X(20, 9, false)         \
  X(10, 4, false)         \
  X(10, 1, false)         \
  X(10, 5, false)

int launch_tangent_mfma(const DevModel& M, const WorkItem* d_work,
                        const std::vector<int>& groups,
                        const std::vector<int>& gstart,
                        const std::vector<int>& gcount, int precision,
                        hipStream_t s) {
  for (size_t k = 0; k < groups.size(); ++k) {
    if (gcount[k] == 0) continue;
    const WorkItem* w = d_work + gstart[k];
    const bool geonl = groups[k] & 1;
    const bool trn = (groups[k] >> 2) & 1;
    const dim3 grid(gcount[k]);
    bool done = false;
#define LAUNCH(NN, G, BB, T, GE, TR)                                          \
  hipLaunchKernelGGL((k_tangent_mfma<NN, G, BB, T, GE, TR>), grid,            \
                     dim3(64 * tan_mfma_waves(NN)), 0, s, M, w,               \
                     precision == W3D_FP32 ? (T*)M.ek32 : (T*)M.ek64)
#define BY_TRN(NN, G, BB, T, GE)          \
  if (trn)                                \
    LAUNCH(NN, G, BB, T, GE, true);       \
  else                                    \
    LAUNCH(NN, G, BB, T, GE, false)
#define X(NN, G, BB)                                                     \
  if (!done && (groups[k] >> 3) == (group_key(NN, G, BB) >> 3)) {       \
    if (precision == W3D_FP32 && geonl) {                                \
      BY_TRN(NN, G, BB, float, true);                                    \
    } else if (precision == W3D_FP32) {                                  \
      BY_TRN(NN, G, BB, float, false);                                   \
    } else if (geonl) {                                                  \
      BY_TRN(NN, G, BB, double, true);                                   \
    } else {                                                             \
      BY_TRN(NN, G, BB, double, false);                                  \
    }                                                                    \
    done = true;                                                         \
  }
    W3D_TAN_GROUPS(X)
#undef X
#undef BY_TRN
#undef LAUNCH
    if (!done) return W3D_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

}  // namespace w3d
