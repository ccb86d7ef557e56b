// Variants of the tile block factor (k_bt_factor in vk_bt.hip) with other
// thread counts per column.
//
// Same algorithm, pivots and per-element arithmetic (bitwise-equal inverse
// stacks and fused forward sweep): the block-Thomas recursion of
// solver.factor_block_thomas_diag_offdiag (solver.py:33-64), each A'_j
// inverted in registers by swap-free Gauss-Jordan with the first-max pivot.
//
// One workgroup of NT threads owns one column: NT = 64 (one wave, no
// barrier between the waves of a column), 128 (two waves) or 512 (eight
// waves, even SL only).  Thread (tr, tc) = (tid/16, tid%16), tr < NT/16,
// holds the R x SL register tile, R = 16 SL / (NT/16),
//   a[r][t] = A[(NT/16) r + tr][16 t + tc]                   (row slot r),
// so each wave's search and fix-ups cover 256/NT times the rows of the
// four-wave kernel.  Entry (i, q) sits where the tile layout of k_bt_factor
// puts it: ((i/16) SL + q/16) * 256 + (i%16) * 16 + q%16.
#include "vk_common.h"

#include <utility>

namespace {

template <int NT>
struct Shape {
  static constexpr int NW = NT / 64;    // waves per column
  static constexpr int TR = NT / VK_G;  // lane rows (tr range)
};

template <int SL, int NT>
struct Bt2Shared {
  static constexpr int NW = Shape<NT>::NW, TR = Shape<NT>::TR;
  static constexpr int XB = (TR > VK_G ? TR : VK_G) * VK_G * SL;
  double cand_row[2][NW][VK_G * SL];  // each wave's best pivot row
  double cand_val[2][NW];
  int cand_idx[2][NW];
  double xbuf[XB];                     // permutation staging (max(TR,16) x N)
  double vec[2][VK_G * SL];            // sweep vector, double-buffered by layer
  int perm[VK_G * SL];
  int iperm[VK_G * SL];
};

// Row of A held in row slot r by lane row tr; row p sits in slot p / TR of
// lane row p % TR.
template <int NT>
__device__ __forceinline__ int row_of(int r, int tr) {
  return Shape<NT>::TR * r + tr;
}

// Offset of entry (r, t) of this thread in one tile-layout block.
template <int SL, int NT>
__device__ __forceinline__ int tile_off(int r, int t, int tid) {
  const int row = row_of<NT>(r, tid >> 4);
  return ((row >> 4) * SL + t) * VK_NT + ((row & 15) << 4) + (tid & 15);
}

template <int SL, int NT>
struct Gj2Ctx {
  Bt2Shared<SL, NT>* sh;
  int n, kb, bad;
  double bias[VK_G * SL / Shape<NT>::TR];  // 0 while the slot's row is a candidate, -inf after
};

// One Gauss-Jordan step with pivot column k = 16 kb + KC (register column 0
// after the column rotation); the arithmetic of gj_step in vk_bt.hip.
template <int SL, int NT, int KC>
__device__ __forceinline__ bool gj2_step(double (&a)[VK_G * SL / Shape<NT>::TR][SL],
                                         Gj2Ctx<SL, NT>& c) {
  constexpr int TR = Shape<NT>::TR, NW = Shape<NT>::NW, R = VK_G * SL / TR;
  const int k = c.kb * VK_G + KC;
  if (k >= c.n) return false;  // uniform: the last block may be partial
  Bt2Shared<SL, NT>& sh = *c.sh;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int buf = k & 1;
  __builtin_amdgcn_s_setprio(2);
  double colv[R], rank[R];
  double best = -INFINITY;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    colv[r] = row_bcast<KC>(a[r][0]);
    rank[r] = fabs(colv[r]) + c.bias[r];
    best = fmax(best, rank[r]);
  }
  const double m = rows_max(best);
  // First row of the wave attaining m: rows grow with the slot r, then tr.
  int brow = 0x7fffffff;
  if (m >= 0.0) {
#pragma unroll
    for (int r = R - 1; r >= 0; --r) {
      const unsigned long long bal = __ballot(rank[r] == m);
      if (bal) brow = row_of<NT>(r, wave * 4 + (int)(__builtin_ctzll(bal) >> 4));
    }
  }
  brow = __builtin_amdgcn_readfirstlane(brow);
  // The owning row group publishes the candidate row (rotated order).
  if (brow != 0x7fffffff) {
    const unsigned wsm = __builtin_amdgcn_readfirstlane(1u << (brow / TR));
    if (tr == brow % TR) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (wsm & (1u << r)) {
#pragma unroll
          for (int t = 0; t < SL; ++t) sh.cand_row[buf][wave][t * VK_G + tc] = a[r][t];
        }
      }
    }
  }
  if (lane == 0) {
    sh.cand_val[buf][wave] = m;
    sh.cand_idx[buf][wave] = brow;
  }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  double pval = sh.cand_val[buf][0];
  int p = sh.cand_idx[buf][0];
  int pw = 0;
#pragma unroll
  for (int w = 1; w < NW; ++w) {
    const double v = sh.cand_val[buf][w];
    const int rr = sh.cand_idx[buf][w];
    if (v > pval || (v == pval && rr < p)) { pval = v; p = rr; pw = w; }
  }
  p = __builtin_amdgcn_readfirstlane(p);
  pw = __builtin_amdgcn_readfirstlane(pw);
  if ((unsigned)p >= (unsigned)c.n) {  // no candidate left: keep indices in range
    p = k;
    c.bad = 1;
  }
  double u[SL];
#pragma unroll
  for (int t = 0; t < SL; ++t) u[t] = sh.cand_row[buf][pw][t * VK_G + tc];
  const double piv = sh.cand_row[buf][pw][KC];
  if (!(piv != 0.0) || !isfinite(piv)) c.bad = 1;
  const double rpiv = 1.0 / piv;
  if (tid == 0) sh.perm[k] = p;
  double l[R];
#pragma unroll
  for (int r = 0; r < R; ++r) l[r] = colv[r] * rpiv;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int t = 0; t < SL; ++t) a[r][t] = fma(-l[r], u[t], a[r][t]);
  if (tc == KC) {
#pragma unroll
    for (int r = 0; r < R; ++r) a[r][0] = -l[r];
  }
  const unsigned psm = __builtin_amdgcn_readfirstlane(1u << (p / TR));
  if (tr == p % TR) {
    double pr[SL];
#pragma unroll
    for (int t = 0; t < SL; ++t) pr[t] = u[t] * rpiv;
    if (tc == KC) pr[0] = rpiv;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (psm & (1u << r)) {
#pragma unroll
        for (int t = 0; t < SL; ++t) a[r][t] = pr[t];
        c.bias[r] = -INFINITY;
      }
    }
  }
  return true;
}

template <int SL, int NT, int... KC>
__device__ __forceinline__ void gj2_block(double (&a)[VK_G * SL / Shape<NT>::TR][SL],
                                          Gj2Ctx<SL, NT>& c, std::integer_sequence<int, KC...>) {
  (void)(gj2_step<SL, NT, KC>(a, c) && ...);
}

// Invert the register-resident matrix in place (natural order on return);
// nonzero in every thread when a zero or non-finite pivot was met.
template <int SL, int NT>
__device__ __forceinline__ int gj2_invert(double (&a)[VK_G * SL / Shape<NT>::TR][SL], int n,
                                          Bt2Shared<SL, NT>& sh) {
  constexpr int TR = Shape<NT>::TR, R = VK_G * SL / TR, N = VK_G * SL;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  Gj2Ctx<SL, NT> c{&sh, n, 0, 0, {}};
#pragma unroll
  for (int r = 0; r < R; ++r) c.bias[r] = (row_of<NT>(r, tr) < n) ? 0.0 : -INFINITY;
#pragma unroll 1
  for (int kb = 0; kb < SL; ++kb) {
    c.kb = kb;
    gj2_block<SL, NT>(a, c, std::make_integer_sequence<int, VK_G>{});
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double first = a[r][0];
#pragma unroll
      for (int t = 0; t + 1 < SL; ++t) a[r][t] = a[r][t + 1];
      a[r][SL - 1] = first;
    }
  }
  const int bad = c.bad;
  for (int i = n + tid; i < N; i += NT) sh.perm[i] = i;
  for (int i = tid; i < N; i += NT) sh.iperm[i] = i;
  __syncthreads();
  for (int i = tid; i < N; i += NT) sh.iperm[sh.perm[i]] = i;
  __syncthreads();
  // R -> X with X[a][b] = R[pi(a)][pi^-1(b)]: the columns one row slot at a
  // time (TR rows x N in xbuf), then the rows one tile column at a time.
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int t = 0; t < SL; ++t) sh.xbuf[tr * N + sh.perm[t * VK_G + tc]] = a[r][t];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < SL; ++t) a[r][t] = sh.xbuf[tr * N + t * VK_G + tc];
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < SL; ++t) {
#pragma unroll
    for (int r = 0; r < R; ++r) sh.xbuf[sh.iperm[row_of<NT>(r, tr)] * VK_G + tc] = a[r][t];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) a[r][t] = sh.xbuf[row_of<NT>(r, tr) * VK_G + tc];
    __syncthreads();
  }
  return bad;
}

// Element (i, q) of a natural-layout block, or of its transpose; zero in
// the padding.
__device__ inline double nat_elem2(const double* blk, int ni, int i, int q, int transpose) {
  if (i >= ni || q >= ni) return 0.0;
  return transpose ? blk[(int64_t)q * ni + i] : blk[(int64_t)i * ni + q];
}

}  // namespace

// k_bt_factor with NT threads per column (see the top of this file): diag
// is natural (TILE_IN false) or tile layout (TILE_IN true, may alias inv);
// a non-null rhs fuses the forward sweep into rmod.  The two-wave
// tile-input kernel is held to 256 VGPRs for SL <= 6 (two waves per SIMD),
// the eight-wave kernel to 128 (two columns per CU); the one-wave kernel may
// take all 512 registers of a lane (VGPRs and AGPRs), one wave per SIMD.
template <int SL, bool TILE_IN, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(
    NT == 512 ? 4 : (NT == 128 && TILE_IN && SL <= 6 ? 2 : 1)))) void k_bt_factor2(
    const VkCol* __restrict__ cols, int ncols, const double* diag, int transpose,
    const double* __restrict__ sup_d, const double* __restrict__ sub_d, double* inv,
    const double* __restrict__ rhs, double* __restrict__ rmod, int* __restrict__ status,
    const int* __restrict__ active) {
  constexpr int R = VK_G * SL / Shape<NT>::TR;
  __shared__ Bt2Shared<SL, NT> sh;
  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const int nz = col.nz, ni = col.ni;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  constexpr int BLK = vk_blk<SL>();
  const double* bsup = sup_d + col.off_s;
  const double* bsub = sub_d + col.off_s;
  double a[R][SL];
  int bad = 0;
  const bool sweep = (rhs != nullptr);
  for (int j = 0; j < nz; ++j) {
    const double* Dj_nat = diag + col.off_s * ni + (int64_t)j * ni * ni;
    const double* Dj_tile = diag + col.off_b + (int64_t)j * BLK;
    if (j == 0) {
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int t = 0; t < SL; ++t) {
          if constexpr (TILE_IN)
            a[r][t] = Dj_tile[tile_off<SL, NT>(r, t, tid)];
          else
            a[r][t] = nat_elem2(Dj_nat, ni, row_of<NT>(r, tr), t * VK_G + tc, transpose);
        }
      if (sweep) {
        for (int i = tid; i < VK_G * SL; i += NT) {
          const double v = (i < ni) ? rhs[col.off_s + i] : 0.0;
          sh.vec[0][i] = v;
          if (i < ni) rmod[col.off_s + i] = v;
        }
      }
    } else {
      double cv[R], bv[SL];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = row_of<NT>(r, tr);
        cv[r] = (i < ni) ? bsub[(int64_t)(j - 1) * ni + i] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < SL; ++t) {
        const int q = t * VK_G + tc;
        bv[t] = (q < ni) ? bsup[(int64_t)(j - 1) * ni + q] : 0.0;
      }
      if (sweep) {
        // r'_j = r_j - c_j o (X_{j-1} r'_{j-1}) while a still holds X_{j-1}.
        const double* v = sh.vec[(j - 1) & 1];
        double vv[SL];
#pragma unroll
        for (int t = 0; t < SL; ++t) vv[t] = v[t * VK_G + tc];
        double w[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          double acc = 0.0;
#pragma unroll
          for (int t = 0; t < SL; ++t) acc = fma(a[r][t], vv[t], acc);
          w[r] = vk_rowsum16(acc);
        }
        if (tc == 0) {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const int i = row_of<NT>(r, tr);
            double rr = 0.0;
            if (i < ni) {
              rr = rhs[col.off_s + (int64_t)j * ni + i] - cv[r] * w[r];
              rmod[col.off_s + (int64_t)j * ni + i] = rr;
            }
            sh.vec[j & 1][i] = rr;
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int t = 0; t < SL; ++t) {
          double d;
          if constexpr (TILE_IN)
            d = Dj_tile[tile_off<SL, NT>(r, t, tid)];
          else
            d = nat_elem2(Dj_nat, ni, row_of<NT>(r, tr), t * VK_G + tc, transpose);
          a[r][t] = d - (cv[r] * bv[t]) * a[r][t];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    bad |= gj2_invert<SL, NT>(a, ni, sh);
    double* dst = inv + col.off_b + (int64_t)j * BLK;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int t = 0; t < SL; ++t) dst[tile_off<SL, NT>(r, t, tid)] = a[r][t];
  }
  if (bad && tid == 0 && status) atomicOr(&status[c], VK_ST_SINGULAR);
}

int vk_launch_bt_factor2(int sl, int nt, const VkCol* cols, int ncols, const double* diag,
                         int diag_tile, int transpose, const double* sup_d, const double* sub_d,
                         double* inv, const double* rhs, double* rmod, int* status,
                         const int* active, hipStream_t stream) {
#define VK_LAUNCH2(n, T, W)                                                             \
  hipLaunchKernelGGL((k_bt_factor2<n, T, W>), dim3(ncols), dim3(W), 0, stream, cols,    \
                     ncols, diag, transpose, sup_d, sub_d, inv, rhs, rmod, status, active)
#define VK_NT2(n, W)                                                  \
  if (nt == W) {                                                      \
    if (diag_tile) VK_LAUNCH2(n, true, W); else VK_LAUNCH2(n, false, W); \
    break;                                                            \
  }
#define VK_CASE2(n)  \
  case n:            \
    VK_NT2(n, 64)    \
    VK_NT2(n, 128)   \
    return VK_ERR_ARG;
#define VK_CASE2E(n) \
  case n:            \
    VK_NT2(n, 64)    \
    VK_NT2(n, 128)   \
    VK_NT2(n, 512)   \
    return VK_ERR_ARG;
  switch (sl) {
    VK_CASE2(1) VK_CASE2E(2) VK_CASE2(3) VK_CASE2E(4) VK_CASE2(5) VK_CASE2E(6) VK_CASE2(7)
    default:
      return VK_ERR_UNSUPPORTED;
  }
#undef VK_CASE2
#undef VK_CASE2E
#undef VK_NT2
#undef VK_LAUNCH2
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
