// w3d_tangent.hip -- block tangent Ke = sum_gp B^T (D w |J|) B on the
// matrix cores (gfx950 MFMA), FP32 (v_mfma_f32_16x16x4f32, FP32
// accumulation) or FP64 (v_mfma_f64_16x16x4f64).
//
// Reference: tanstf.f:119 do_nlek_block, rktstf.f:15, gptns1.f:16 gptns1,
// gptns1.f:1174 extract_D_symmetric, gptns1.f:1628 bdbtgen, gpmas1.f:399
// adjust_cnst.  bdbtgen forms, per point, BD = D B and then the upper
// triangle Ke(row,col) += sum_r B(r,col) BD(r,row).  Here the sum over
// points and strain components is one contraction over k = (point, r):
//   Ke = A B   with  A(row, k) = BD(r, row) at point p,  B(k, col) = B(r, col)
// taken 4 k at a time by the 16x16x4 MFMA on 16x16 tiles of the upper
// triangle.  One wave owns one element: per chunk of two points its lanes
// build the B and BD columns in FP64 (one (point, dof) pair per lane) from
// the element geometry and the stored [D], round them to the MFMA type into
// a wave-private LDS slab, and then feed the 16x16x4 steps from it.
#include <hip/hip_runtime.h>

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

template <int NN, int G, bool BBAR, typename T>
__global__ void __launch_bounds__(64 * tan_mfma_waves(NN))
    k_tangent_mfma(DevModel M, const WorkItem* __restrict__ work,
                   T* __restrict__ ek_out) {
  constexpr int W = tan_mfma_waves(NN);   // waves per workgroup
  constexpr int TD = 3 * NN;              // element dofs
  constexpr int NTL = (TD + 15) / 16;     // 16-wide tiles per side
  constexpr int TP = 16 * NTL;            // padded dof count
  constexpr int NUP = NTL * (NTL + 1) / 2;  // upper tiles
  constexpr int NCH = (G + 1) / 2;        // chunks of two points (12 k)
  constexpr int LDP = TP + 16;            // padded LDS row
  constexpr int ND = (21 * G + 63) / 64;  // [D] words per lane
  using MF = Mfma<T>;
  using v4 = typename MF::v4;

  __shared__ double s_dn[3][G][NN];
  __shared__ double s_w[G];
  __shared__ double s_ce[W][TD];
  __shared__ double s_gam[W][G][10];
  __shared__ double s_vol[W][24];
  __shared__ double s_d[W][G][21];
  __shared__ int s_node[W][NN];
  __shared__ int s_inv[W][NN];
  // per wave: B(r, dof) and BD(r, dof) = (D B)(r, dof) of the chunk's two
  // points; after the contraction the same slab stages the packed Ke
  __shared__ T s_bd[W][2][12][LDP];
  static_assert(24 * LDP >= TD * (TD + 1) / 2, "Ke staging fits the slab");

  const WorkItem wi = work[blockIdx.x];
  const BlockDesc bd = M.blocks[wi.blk];
  const GaussTable* gt = M.tables + bd.table;
  const int tid = threadIdx.x;
  const int wv = tid >> 6;   // wave of this lane
  const int l = tid & 63;    // lane in the wave
  const int q = l >> 4;      // k row inside a 4-wide k step
  const int cres = l & 15;   // column residue
  const double scalar = bd.etype == W3D_TET10 ? 1.0 / 6.0 : 1.0;
  const double third = 1.0 / 3.0;
  const double alpha = (2.0 + M.eps_bbar) * third;
  const double beta_b = (1.0 - M.eps_bbar) * third;
  const int64_t SG = (int64_t)bd.span * G;
  const double* cep_b = M.cep + 21 * bd.q_off;

  for (int k = tid; k < 3 * G * NN; k += 64 * W) {
    int d = k / (G * NN), rem = k % (G * NN);
    s_dn[d][rem / NN][rem % NN] = gt->dn[d][rem / NN][rem % NN];
  }
  if (tid < G) s_w[tid] = gt->w[tid];
  __syncthreads();

  // element data in flight: coordinates of node l (lanes < NN) and this
  // lane's share of the stored [D]
  double pc[3], pd[ND];
  int pnode = 0;
  auto fetch = [&](int e) {
    const int ie = wi.first + e;
    if (l < NN) {
      int node = M.conn[bd.conn_off + (int64_t)ie * NN + l];
      pnode = node;
      pc[0] = M.coords[3 * node];
      pc[1] = M.coords[3 * node + 1];
      pc[2] = M.coords[3 * node + 2];
    }
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      int k = l + 64 * i;
      if (k < 21 * G) pd[i] = cep_b[(k / G) * SG + (int64_t)ie * G + k % G];
    }
  };
  auto stage = [&]() {
    if (l < NN) {
      s_node[wv][l] = pnode;
      s_ce[wv][l] = pc[0];
      s_ce[wv][NN + l] = pc[1];
      s_ce[wv][2 * NN + l] = pc[2];
    }
#pragma unroll
    for (int i = 0; i < ND; ++i) {
      int k = l + 64 * i;
      if (k < 21 * G) s_d[wv][k % G][k / G] = pd[i];
    }
  };

  int e = wv;
  if (e < wi.count) {
    fetch(e);
    stage();
  }
  for (; e < wi.count; e += W) {
    const int ie = wi.first + e;
    wave_sync();
    // P1: Jacobian at every point (jacob1)
    if (l < G) {
      const int g = l;
      double j11 = 0, j12 = 0, j13 = 0, j21 = 0, j22 = 0, j23 = 0, j31 = 0,
             j32 = 0, j33 = 0;
      for (int j = 0; j < NN; ++j) {
        double x = s_ce[wv][j], y = s_ce[wv][NN + j], z = s_ce[wv][2 * NN + j];
        double a = s_dn[0][g][j], b = s_dn[1][g][j], c = s_dn[2][g][j];
        j11 = j11 + a * x; j12 = j12 + a * y; j13 = j13 + a * z;
        j21 = j21 + b * x; j22 = j22 + b * y; j23 = j23 + b * z;
        j31 = j31 + c * x; j32 = j32 + c * y; j33 = j33 + c * z;
      }
      double c1 = j22 * j33 - j23 * j32;
      double c2 = j21 * j33 - j23 * j31;
      double c3 = j21 * j32 - j22 * j31;
      double dj = j11 * c1 - j12 * c2 + j13 * c3;
      double* ga = s_gam[wv][g];
      ga[0] = c1 / dj;
      ga[3] = -c2 / dj;
      ga[6] = c3 / dj;
      ga[1] = (j32 * j13 - j12 * j33) / dj;
      ga[4] = (j11 * j33 - j31 * j13) / dj;
      ga[7] = (j12 * j31 - j11 * j32) / dj;
      ga[2] = (j12 * j23 - j13 * j22) / dj;
      ga[5] = (j13 * j21 - j11 * j23) / dj;
      ga[8] = (j11 * j22 - j12 * j21) / dj;
      ga[9] = dj;
    }
    wave_sync();
    // P2: B-bar volume terms; [D] scaled by w |J| (extract_D_symmetric)
    // and by 1/6 for tetrahedra (adjust_cnst)
    if (BBAR && l < 8) {
      const int j = l;
      double vx = 0, vy = 0, vz = 0, volume = 0;
      for (int g = 0; g < G; ++g) {
        const double* a = s_gam[wv][g];
        double d = a[9];
        double nx = s_dn[0][g][j], ny = s_dn[1][g][j], nz = s_dn[2][g][j];
        vx = vx + (a[0] * nx + a[1] * ny + a[2] * nz) * d;
        vy = vy + (a[3] * nx + a[4] * ny + a[5] * nz) * d;
        vz = vz + (a[6] * nx + a[7] * ny + a[8] * nz) * d;
        volume = volume + d;
      }
      double voli = 1.0 / volume;
      s_vol[wv][3 * j] = vx * voli;
      s_vol[wv][3 * j + 1] = vy * voli;
      s_vol[wv][3 * j + 2] = vz * voli;
    }
    for (int k = l; k < 21 * G; k += 64) {
      int g = k / 21, c = k % 21;
      double v = (s_w[g] * s_gam[wv][g][9]) * s_d[wv][g][c];
      if (bd.etype == W3D_TET10) v = scalar * v;
      s_d[wv][g][c] = v;
    }
    // the next element's data travels while this one is contracted
    const bool more = e + W < wi.count;
    if (more) fetch(e + W);
    wave_sync();
    // P3: the contraction on the matrix cores, two points per chunk
    v4 acc[NUP];
#pragma unroll
    for (int t = 0; t < NUP; ++t) acc[t] = v4{0, 0, 0, 0};
    for (int ch = 0; ch < NCH; ++ch) {
      // (a) B and BD columns of (point, dof) pairs into the wave's slab
      for (int k = l; k < 2 * TP; k += 64) {
        const int p = k / TP, dof = k % TP, g = 2 * ch + p;
        double col[6] = {0, 0, 0, 0, 0, 0};
        if (g < G && dof < TD) {
          const int c = dof / NN, j = dof % NN;
          const double* ga = s_gam[wv][g];
          double nx = s_dn[0][g][j], ny = s_dn[1][g][j], nz = s_dn[2][g][j];
          double bx = ga[0] * nx + ga[1] * ny + ga[2] * nz;
          double by = ga[3] * nx + ga[4] * ny + ga[5] * nz;
          double bz = ga[6] * nx + ga[7] * ny + ga[8] * nz;
          if (c == 0) {
            col[0] = bx; col[3] = by; col[5] = bz;
            if (BBAR) {
              double vx = s_vol[wv][3 * j];
              col[0] = alpha * bx + beta_b * vx;
              col[1] = col[2] = beta_b * (vx - bx);
            }
          } else if (c == 1) {
            col[1] = by; col[3] = bx; col[4] = bz;
            if (BBAR) {
              double vy = s_vol[wv][3 * j + 1];
              col[0] = col[2] = beta_b * (vy - by);
              col[1] = alpha * by + beta_b * vy;
            }
          } else {
            col[2] = bz; col[4] = by; col[5] = bx;
            if (BBAR) {
              double vz = s_vol[wv][3 * j + 2];
              col[0] = col[1] = beta_b * (vz - bz);
              col[2] = alpha * bz + beta_b * vz;
            }
          }
        }
        const double* dm = s_d[wv][g < G ? g : 0];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
          // BD(r, dof) = sum_m D(m, r) B(m, dof), bdbtgen order
          double db = dm[tri(0, r)] * col[0] + dm[tri(1, r)] * col[1] +
                      dm[tri(2, r)] * col[2] + dm[tri(3, r)] * col[3] +
                      dm[tri(4, r)] * col[4] + dm[tri(5, r)] * col[5];
          s_bd[wv][0][6 * p + r][dof] = (T)col[r];
          s_bd[wv][1][6 * p + r][dof] = (T)db;
        }
      }
      wave_sync();
      // (b) three k steps of 4 rows: A from BD (row dofs), B from B (col)
#pragma unroll
      for (int ks = 0; ks < 3; ++ks) {
        const int kr = 4 * ks + q;
        T av[NTL], bv[NTL];
#pragma unroll
        for (int t = 0; t < NTL; ++t) {
          av[t] = s_bd[wv][1][kr][16 * t + cres];
          bv[t] = s_bd[wv][0][kr][16 * t + cres];
        }
#pragma unroll
        for (int t = 0; t < NUP; ++t)
          acc[t] = MF::mma(av[tile_i(t)], bv[tile_j(t)], acc[t]);
      }
      wave_sync();
    }
    // P4: the upper triangle staged packed (column-major) in the wave's
    // slab, then stored contiguously in the node-pair block order
    // (npb_pos): word k gathers Ke(p, q) of the nodes of ranks r and s
    if (l < NN) {
      int r = 0;
      const int mine = s_node[wv][l];
#pragma unroll
      for (int b = 0; b < NN; ++b) r += s_node[wv][b] < mine ? 1 : 0;
      s_inv[wv][r] = l;
    }
    T* stg = &s_bd[wv][0][0][0];
#pragma unroll
    for (int t = 0; t < NUP; ++t) {
      const int I = tile_i(t), J = tile_j(t);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * I + MF::row(l, v), col = 16 * J + cres;
        if (row <= col && col < TD) stg[col * (col + 1) / 2 + row] = acc[t][v];
      }
    }
    wave_sync();
    T* out = ek_out + bd.ek_off + (int64_t)ie * (TD * (TD + 1) / 2);
    const uint16_t* tab = M.npb_tab + npb_tab_off(NN);
    for (int k = l; k < TD * (TD + 1) / 2; k += 64) {
      const int code = tab[k];
      const int a = s_inv[wv][code & 31], b = s_inv[wv][(code >> 5) & 31];
      int p = ((code >> 10) & 3) * NN + a, q = (code >> 12) * NN + b;
      if (p > q) {
        int x = p;
        p = q;
        q = x;
      }
      out[k] = stg[q * (q + 1) / 2 + p];
    }
    wave_sync();
    if (more) stage();
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
    bool done = false;
#define X(NN, G, BB)                                                       \
  if (!done && groups[k] == group_key(NN, G, BB)) {                         \
    if (precision == W3D_FP32)                                             \
      hipLaunchKernelGGL((k_tangent_mfma<NN, G, BB, float>),               \
                         dim3(gcount[k]), dim3(64 * tan_mfma_waves(NN)), 0, \
                         s, M, w, M.ek32);                                 \
    else                                                                   \
      hipLaunchKernelGGL((k_tangent_mfma<NN, G, BB, double>),              \
                         dim3(gcount[k]), dim3(64 * tan_mfma_waves(NN)), 0, \
                         s, M, w, M.ek64);                                 \
    done = true;                                                           \
  }
    W3D_TAN_GROUPS(X)
#undef X
    if (!done) return W3D_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? W3D_OK : W3D_ERR_DEVICE;
}

}  // namespace w3d
