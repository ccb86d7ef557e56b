// Block-tridiagonal Thomas factor and solves for diagonal off-diagonal blocks.
//
// Restates solver.factor_block_thomas_diag_offdiag (solver.py:33-64) and
// solver.solve_block_thomas_diag_offdiag (solver.py:67-112):
//   A'_0 = D_0,  A'_j = D_j - (c_j b_{j-1}^T) o inv(A'_{j-1})
//   r'_0 = r_0,  r'_j = r_j - c_j o inv(A'_{j-1}) r'_{j-1}
//   x_last = inv(A'_last) r'_last,  x_j = inv(A'_j) (r'_j - b_j o x_{j+1})
// with b = sup_d, c = sub_d.  The reference keeps LU factors and calls
// lu_solve(., I) for the inverse every layer; here each A'_j is inverted
// once, in registers, by Gauss-Jordan elimination with partial pivoting (the
// same pivot sequence as getrf's), and the stack of inverses replaces the
// LU stack, so both sweeps become matrix-vector products.
//
// One workgroup of 256 threads owns one column.  Thread (tr, tc) = (tid/16,
// tid%16) holds the SL x SL register tile a[s][t] = A[16s+tr][16t+tc].
#include "vk_common.h"

#include <utility>

namespace {

constexpr int NW = VK_NT / 64;  // waves per workgroup

template <int SL>
struct BtShared {
  double cand_row[2][NW][VK_G * SL];  // each wave's best pivot row
  double cand_val[2][NW];
  int cand_idx[2][NW];
  double xbuf[VK_G * VK_G * SL];      // permutation staging (16 x N)
  double vec[2][VK_G * SL];           // sweep vector, double-buffered by layer
  int perm[VK_G * SL];
  int iperm[VK_G * SL];
};

template <int SL>
struct GjCtx {
  BtShared<SL>* sh;
  int n, kb, bad;
  // per tile row s: 0 while row 16s+tr is a row of A not yet used as a
  // pivot, -inf after (and for padding rows), added to |A[i][k]| to rank it
  double bias[SL];
};

// One Gauss-Jordan step with pivot column k = 16 kb + KC, which the column
// rotation keeps in register column 0.  Returns false once k reaches n.
//
// The pivot is the first maximum of |A[i][k]| over the unused rows (the
// getrf sequence).  A row's rank is |A[i][k]| + bias: used rows rank -inf
// (or NaN, for an infinite entry), and fmax passes over NaN, so NaN entries
// are never chosen; a step left without a candidate flags the matrix.  Per-tile-row branches test bits of a mask
// that passes through readfirstlane, so the compiler keeps them as separate
// uniform branches instead of merging the register-tile stores into one
// dynamically indexed store (which would move the tile to scratch memory).
template <int SL, int KC>
__device__ __forceinline__ bool gj_step(double (&a)[SL][SL], GjCtx<SL>& c) {
  const int k = c.kb * VK_G + KC;
  if (k >= c.n) return false;  // uniform: the last block may be partial
  BtShared<SL>& sh = *c.sh;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int buf = k & 1;
  // The search and publication run at raised wave priority: the SIMDs are
  // shared with other columns' waves, and a column's barrier waits for its
  // slowest wave.
  __builtin_amdgcn_s_setprio(2);
  // Column k of this thread's rows and their pivot ranks.
  double colv[SL], rank[SL];
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    colv[s] = row_bcast<KC>(a[s][0]);
    rank[s] = fabs(colv[s]) + c.bias[s];
  }
  // the largest rank by a pairwise tree (fmax is exact and passes over
  // NaN in any order, so the tree gives the chain's value)
  double tmax[SL];
#pragma unroll
  for (int s = 0; s < SL; ++s) tmax[s] = rank[s];
#pragma unroll
  for (int h = 1; h < SL; h *= 2)
#pragma unroll
    for (int s = 0; s + h < SL; s += 2 * h) tmax[s] = fmax(tmax[s], tmax[s + h]);
  const double best = fmax(-INFINITY, tmax[0]);
  const double m = rows_max(best);
  // First row of the wave attaining m: smallest tile row s, then smallest tr.
  int brow = 0x7fffffff;
  if (m >= 0.0) {
#pragma unroll
    for (int s = SL - 1; s >= 0; --s) {
      const unsigned long long bal = __ballot(rank[s] == m);
      if (bal) brow = s * VK_G + wave * 4 + (int)(__builtin_ctzll(bal) >> 4);
    }
  }
  brow = __builtin_amdgcn_readfirstlane(brow);
  // The owning row group publishes the whole candidate row (rotated order).
  if (brow != 0x7fffffff) {
    const unsigned wsm = __builtin_amdgcn_readfirstlane(1u << (brow >> 4));
    if (tr == (brow & 15)) {
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        if (wsm & (1u << s)) {
#pragma unroll
          for (int t = 0; t < SL; ++t) sh.cand_row[buf][wave][t * VK_G + tc] = a[s][t];
        }
      }
    }
  }
  if (lane == 0) { sh.cand_val[buf][wave] = m; sh.cand_idx[buf][wave] = brow; }
  __builtin_amdgcn_s_setprio(0);
  __syncthreads();
  // Global pivot: best of the wave candidates.  All candidates are read
  // first and combined with selects, so the LDS reads are one round trip.
  double cv[NW];
  int ci[NW];
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    cv[w] = sh.cand_val[buf][w];
    ci[w] = sh.cand_idx[buf][w];
  }
  double pval = cv[0];
  int p = ci[0];
#pragma unroll
  for (int w = 1; w < NW; ++w) {
    const bool take = cv[w] > pval || (cv[w] == pval && ci[w] < p);
    pval = take ? cv[w] : pval;
    p = take ? ci[w] : p;
  }
  p = __builtin_amdgcn_readfirstlane(p);
  if ((unsigned)p >= (unsigned)c.n) {  // no candidate left: keep indices in range
    p = k;
    c.bad = 1;
  }
  // row 16 s + tr lives in wave tr / 4, which published it
  const int pw = (p >> 2) & (NW - 1);
  // the pivot first, so its reciprocal can start while the row arrives
  const double piv = sh.cand_row[buf][pw][KC];
  double u[SL];
#pragma unroll
  for (int t = 0; t < SL; ++t) u[t] = sh.cand_row[buf][pw][t * VK_G + tc];
  if (!(piv != 0.0) || !isfinite(piv)) c.bad = 1;
  const double rpiv = 1.0 / piv;
  if (tid == 0) sh.perm[k] = p;
  // Rank-1 update of every row; column k and row p are rewritten below.
  double l[SL];
#pragma unroll
  for (int s = 0; s < SL; ++s) l[s] = colv[s] * rpiv;
#pragma unroll
  for (int s = 0; s < SL; ++s)
#pragma unroll
    for (int t = 0; t < SL; ++t) a[s][t] = fma(-l[s], u[t], a[s][t]);
  if (tc == KC) {
#pragma unroll
    for (int s = 0; s < SL; ++s) a[s][0] = -l[s];
  }
  // Row p becomes u / piv, with 1 / piv in column k; the products are
  // formed in the branch of p's tile row, straight into the tile.
  const unsigned psm = __builtin_amdgcn_readfirstlane(1u << (p >> 4));
  if (tr == (p & 15)) {
    const double u0 = tc == KC ? 1.0 : u[0];
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      if (psm & (1u << s)) {
        a[s][0] = u0 * rpiv;
#pragma unroll
        for (int t = 1; t < SL; ++t) a[s][t] = u[t] * rpiv;
        c.bias[s] = -INFINITY;
      }
    }
  }
  return true;
}

// The 16 steps of column block kb, in order, stopping at n.
template <int SL, int... KC>
__device__ __forceinline__ void gj_block(double (&a)[SL][SL], GjCtx<SL>& c,
                                         std::integer_sequence<int, KC...>) {
  (void)(gj_step<SL, KC>(a, c) && ...);
}

// Invert the register-resident matrix a (n x n inside the padded N x N tile)
// in place.  On return a holds inv(A) in natural order.  Returns nonzero
// in every thread when a zero or non-finite pivot was met.
//
// The tile columns rotate left by one after every 16-column block, so the
// pivot column of step k = 16 kb + kc is register column 0 with kc static,
// and the block loop stays rolled; after SL blocks the order is the
// original one again.
template <int SL>
__device__ int gj_invert(double (&a)[SL][SL], int n, BtShared<SL>& sh) {
  const int tid = threadIdx.x;
  const int tr = tid >> 4, tc = tid & 15;
  GjCtx<SL> c{&sh, n, 0, 0, {}};
#pragma unroll
  for (int s = 0; s < SL; ++s) c.bias[s] = (s * VK_G + tr < n) ? 0.0 : -INFINITY;
#pragma unroll 1
  for (int kb = 0; kb < SL; ++kb) {
    c.kb = kb;
    gj_block<SL>(a, c, std::make_integer_sequence<int, VK_G>{});
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const double first = a[s][0];
#pragma unroll
      for (int t = 0; t + 1 < SL; ++t) a[s][t] = a[s][t + 1];
      a[s][SL - 1] = first;
    }
  }
  const int bad = c.bad;
  // Padding indices map to themselves.
  for (int i = n + tid; i < VK_G * SL; i += VK_NT) sh.perm[i] = i;
  for (int i = tid; i < VK_G * SL; i += VK_NT) sh.iperm[i] = i;
  __syncthreads();
  for (int i = tid; i < VK_G * SL; i += VK_NT) sh.iperm[sh.perm[i]] = i;
  __syncthreads();
  // R -> X with X[a][b] = R[pi(a)][pi^-1(b)]: first the columns
  // (R'[r][pi(s)] = R[r][s]), then the rows (X[pi^-1(r)][.] = R'[r][.]).
  constexpr int N = VK_G * SL;
#pragma unroll
  for (int s = 0; s < SL; ++s) {
#pragma unroll
    for (int t = 0; t < SL; ++t) sh.xbuf[tr * N + sh.perm[t * VK_G + tc]] = a[s][t];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < SL; ++t) a[s][t] = sh.xbuf[tr * N + t * VK_G + tc];
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < SL; ++t) {
#pragma unroll
    for (int s = 0; s < SL; ++s) sh.xbuf[sh.iperm[s * VK_G + tr] * VK_G + tc] = a[s][t];
    __syncthreads();
#pragma unroll
    for (int s = 0; s < SL; ++s) a[s][t] = sh.xbuf[(s * VK_G + tr) * VK_G + tc];
    __syncthreads();
  }
  return bad;
}

// y_i = sum_q a[i][q] v[q] for this thread's rows, v in shared memory.
template <int SL>
__device__ inline void tile_matvec(const double (&a)[SL][SL], const double* v, double (&y)[SL]) {
  const int tc = threadIdx.x & 15;
  double vv[SL];
#pragma unroll
  for (int t = 0; t < SL; ++t) vv[t] = v[t * VK_G + tc];
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < SL; ++t) acc = fma(a[s][t], vv[t], acc);
    y[s] = vk_rowsum16(acc);
  }
}

// y_i = sum_q B[i][q] v[q] for this thread's rows, B a tile-layout block
// read straight from memory (not kept in registers), v in shared memory.
template <int SL>
__device__ inline void stream_matvec(const double* __restrict__ src, const double* v,
                                     double (&y)[SL]) {
  const int tid = threadIdx.x, tc = tid & 15;
  double vv[SL];
#pragma unroll
  for (int t = 0; t < SL; ++t) vv[t] = v[t * VK_G + tc];
#pragma unroll
  for (int s = 0; s < SL; ++s) {
    double acc = 0.0;
#pragma unroll
    for (int t = 0; t < SL; ++t) acc = fma(src[(s * SL + t) * VK_NT + tid], vv[t], acc);
    y[s] = vk_rowsum16(acc);
  }
}

template <int SL>
__device__ inline void load_tile(double (&a)[SL][SL], const double* src) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int s = 0; s < SL; ++s)
#pragma unroll
    for (int t = 0; t < SL; ++t) a[s][t] = src[(s * SL + t) * VK_NT + tid];
}

template <int SL>
__device__ inline void store_tile(const double (&a)[SL][SL], double* dst) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int s = 0; s < SL; ++s)
#pragma unroll
    for (int t = 0; t < SL; ++t) dst[(s * SL + t) * VK_NT + tid] = a[s][t];
}

// Element (i, q) of block j of a natural-layout (nz, ni, ni) stack, or of
// its transpose; zero in the padding.
__device__ inline double nat_elem(const double* blk, int ni, int i, int q, int transpose) {
  if (i >= ni || q >= ni) return 0.0;
  return transpose ? blk[(int64_t)q * ni + i] : blk[(int64_t)i * ni + q];
}

}  // namespace

// Factor one column per workgroup.  diag is either natural (nz, ni, ni)
// blocks at off_s*ni (TILE_IN false) or tile-layout blocks at off_b
// (TILE_IN true, which may alias inv).  When rhs is non-null the forward
// sweep r' is fused and written to rmod.  The two input layouts are separate
// instantiations: the natural path's bounds and transposition arithmetic
// would otherwise stay live beside the register tile.  The tile-input kernel
// is held to 128 VGPRs for SL <= 6 (four columns per CU); what it spills is
// per-layer state, outside the pivot steps.  When schur is non-null (tile
// layout at off_b) each Schur complement A'_j is stored there before its
// inversion, for the refined solves.
template <int SL, bool TILE_IN>
__global__ __launch_bounds__(VK_NT) __attribute__((amdgpu_waves_per_eu(TILE_IN && SL <= 6 ? 4 : 1))) void k_bt_factor(
    const VkCol* __restrict__ cols, int ncols, const double* diag,
    int transpose, const double* __restrict__ sup_d, const double* __restrict__ sub_d,
    double* inv, const double* __restrict__ rhs, double* __restrict__ rmod,
    int* __restrict__ status, const int* __restrict__ active, double* __restrict__ schur) {
  __shared__ BtShared<SL> sh;
  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const int nz = col.nz, ni = col.ni;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  constexpr int BLK = vk_blk<SL>();
  const double* bsup = sup_d + col.off_s;  // (nz-1, ni)
  const double* bsub = sub_d + col.off_s;
  double a[SL][SL];
  int bad = 0;
  const bool sweep = (rhs != nullptr);
  for (int j = 0; j < nz; ++j) {
    const double* Dj_nat = diag + col.off_s * ni + (int64_t)j * ni * ni;
    const double* Dj_tile = diag + col.off_b + (int64_t)j * BLK;
    if (j == 0) {
      if constexpr (TILE_IN) {
        load_tile<SL>(a, Dj_tile);
      } else {
#pragma unroll
        for (int s = 0; s < SL; ++s)
#pragma unroll
          for (int t = 0; t < SL; ++t)
            a[s][t] = nat_elem(Dj_nat, ni, s * VK_G + tr, t * VK_G + tc, transpose);
      }
      if (sweep) {
        for (int i = tid; i < VK_G * SL; i += VK_NT) {
          const double v = (i < ni) ? rhs[col.off_s + i] : 0.0;
          sh.vec[0][i] = v;
          if (i < ni) rmod[col.off_s + i] = v;
        }
      }
    } else {
      double cv[SL], bv[SL];
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int i = s * VK_G + tr;
        cv[s] = (i < ni) ? bsub[(int64_t)(j - 1) * ni + i] : 0.0;
      }
#pragma unroll
      for (int t = 0; t < SL; ++t) {
        const int q = t * VK_G + tc;
        bv[t] = (q < ni) ? bsup[(int64_t)(j - 1) * ni + q] : 0.0;
      }
      if (sweep) {
        // r'_j = r_j - c_j o (X_{j-1} r'_{j-1}) while a still holds X_{j-1}.
        double w[SL];
        tile_matvec<SL>(a, sh.vec[(j - 1) & 1], w);
        if (tc == 0) {
#pragma unroll
          for (int s = 0; s < SL; ++s) {
            const int i = s * VK_G + tr;
            double r = 0.0;
            if (i < ni) {
              r = rhs[col.off_s + (int64_t)j * ni + i] - cv[s] * w[s];
              rmod[col.off_s + (int64_t)j * ni + i] = r;
            }
            sh.vec[j & 1][i] = r;
          }
        }
      }
      // One tile row of D_j at a time, so the loads do not pile up beside
      // the register tile.
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int s = 0; s < SL; ++s) {
#pragma unroll
        for (int t = 0; t < SL; ++t) {
          double d;
          if constexpr (TILE_IN)
            d = Dj_tile[(s * SL + t) * VK_NT + tid];
          else
            d = nat_elem(Dj_nat, ni, s * VK_G + tr, t * VK_G + tc, transpose);
          a[s][t] = d - (cv[s] * bv[t]) * a[s][t];
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (schur) store_tile<SL>(a, schur + col.off_b + (int64_t)j * BLK);
    bad |= gj_invert<SL>(a, ni, sh);
    store_tile<SL>(a, inv + col.off_b + (int64_t)j * BLK);
  }
  if (bad && tid == 0 && status) atomicOr(&status[c], VK_ST_SINGULAR);
}

// Sweeps with the inverse stack.  mode 0: forward + backward from rhs;
// mode 1: backward only, rhs already holds r'.  Optional epilogues per
// element of x_j (all pointers at off_s):
//   epi 1: out1 = y + x / e0                          (Ros2 stage-1 yk2)
//   epi 2: out1 = y + e0*k1 + e1*x, out2 = |out1 - yk2|  (Ros2 sol, delta)
//
// REFINE 1 (dstk holds the tile-layout diagonal blocks D_j the inverses were
// formed from): each block solve of the backward sweep takes one step of
// iterative refinement,
//   x_j = X_j t_j + X_j (t_j - A'_j X_j t_j),
//   A'_j v = D_j v - c_{j-1} o (X_{j-1} (b_{j-1} o v))   (b = sup, c = sub),
// A'_j being the Schur complement whose inverse X_j is.  A product with an
// explicit inverse is not backward stable; the refined block solve is, as
// the reference's lu_solve on A'_j is (solver.py:91-112).  The correction of
// x_j is made in the next step down, where X_{j-1} is loaded anyway and X_j
// is still in registers, so a refined sweep reads each inverse once and D
// once.
//
// REFINE 2 (dstk holds the Schur complements A'_j themselves, stored by the
// factor): the same refinement of each block solve of the backward sweep,
// with the residual t_j - A'_j x0_j taken on the stored A'_j, so a layer
// reads X_j and A'_j and needs no second inverse in registers.
//
// refl (REFINE 1 and 2, optional): per column, 0 when the block refinement's
// correction of some layer exceeded that layer's solution in max norm (an
// inverse X_j with no correct digit there, cond(A'_j) ~ 1e12 and up), else
// VK_ST_DONE.  The host gives those columns one whole-column refinement step
// (k_bt_residual, then a correction solve with epi.base) and passes refl as
// the active mask of both, so the other columns skip them.

// The epilogue with its inputs already read: bv = base[e], yv = y[e],
// kv = k1[e], zv = yk2[e] (those the epilogue kind uses).
__device__ inline void solve_epilogue_v(const SolveEpi& epi, int64_t e, double xv, double* x,
                                        double bv, double yv, double kv, double zv) {
  if (epi.base) xv = bv + xv;
  x[e] = xv;
  if (epi.kind == 1) {
    epi.out1[e] = yv + xv / epi.e0;
  } else if (epi.kind == 2) {
#pragma clang fp contract(off)
    const double sol = yv + epi.e0 * kv + epi.e1 * xv;
    epi.out1[e] = sol;
    epi.out2[e] = fabs(sol - zv);
  }
}

__device__ inline void solve_epilogue(const SolveEpi& epi, int64_t e, double xv, double* x) {
  solve_epilogue_v(epi, e, xv, x, epi.base ? epi.base[e] : 0.0, epi.kind ? epi.y[e] : 0.0,
                   epi.kind == 2 ? epi.k1[e] : 0.0, epi.kind == 2 ? epi.yk2[e] : 0.0);
}

// Read-ahead of the per-row vectors a layer's sweep step reads besides its
// block (right-hand sides, off-diagonals, epilogue inputs): NV vectors of
// VK_G * SL rows, spread over the workgroup's threads, loaded together with
// the layer's block so their latency hides behind it, then put in LDS
// (slot v, row i at dst[v * VK_G * SL + i]) for the step that needs them.
// src[v] points at row 0 of vector v in that layer, or is null (zeros).
template <int SL, int NV>
struct RowAhead {
  static constexpr int N = VK_G * SL, PER = (NV * N + VK_NT - 1) / VK_NT;
  double v[PER];
  __device__ __forceinline__ void load(const double* const (&src)[NV], int ni) {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int f = threadIdx.x + q * VK_NT;
      const int w = f / N, i = f - w * N;
      const double* p = src[0];
#pragma unroll
      for (int u = 1; u < NV; ++u) p = w == u ? src[u] : p;
      v[q] = (f < NV * N && p && i < ni) ? p[i] : 0.0;
    }
  }
  __device__ __forceinline__ void store(double* dst) const {
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int f = threadIdx.x + q * VK_NT;
      if (f < NV * N) dst[f] = v[q];
    }
  }
};

// Waves per SIMD the sweep kernel is compiled for.  REFINE 3 holds two and
// a half blocks in registers (X_j across the layer, A'_j beside it, part of
// X_{j-1} read ahead): four columns per CU up to SL 3, three at SL 4, two
// at SL 5-6.
constexpr int solve_waves(int sl, int refine) {
  return sl > 6 ? 1 : refine == 1 ? 2 : refine == 3 ? (sl <= 3 ? 4 : sl == 4 ? 3 : 2) : 4;
}

// REFINE 3 is REFINE 2 with X_j loaded once per layer and kept in
// registers for both of its products (two block reads per refined layer
// instead of three), at fewer columns per CU; the forward sweep reads each
// block one layer ahead.  The arithmetic is REFINE 2's, so the two give the
// same bits.
template <int SL, int REFINE>
__global__ __launch_bounds__(VK_NT) __attribute__((amdgpu_waves_per_eu(solve_waves(SL, REFINE)))) void k_bt_solve(
    const VkCol* __restrict__ cols, int ncols, const double* __restrict__ inv,
    const double* __restrict__ sup_d, const double* __restrict__ sub_d,
    const double* __restrict__ rhs, double* __restrict__ rscratch, double* x,
    int mode, SolveEpi epi, const int* __restrict__ active, const double* __restrict__ dstk,
    int* __restrict__ refl) {
  __shared__ double vec[2][VK_G * SL];
  __shared__ double rv[VK_G * SL];  // refinement: x_j, then the residual
  __shared__ double bx[VK_G * SL];  // refinement: b_{j-1} o x_j
  __shared__ double red[VK_NT / 64][2];  // per wave: max |correction|, max |x| of a layer
  // read-ahead rows, double-buffered by layer: the forward sweep's r and
  // c_{j-1}; the refined backward sweep's base, y, k1, yk2, r'_{j-1}, b_{j-1}
  __shared__ double ahead[2][6 * VK_G * SL];
  constexpr int NR = VK_G * SL;
  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) {
    if (refl && threadIdx.x == 0) refl[c] = VK_ST_DONE;
    return;
  }
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const int nz = col.nz, ni = col.ni;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  constexpr int BLK = vk_blk<SL>();
  const double* bsup = sup_d + col.off_s;
  const double* bsub = sub_d + col.off_s;
  const double* stk = inv + col.off_b;
  double a[SL][SL];
  double w[SL];
  const double* rm = rhs;  // r' source for the backward sweep
  if (mode == 0) {
    for (int i = tid; i < VK_G * SL; i += VK_NT) {
      const double v = (i < ni) ? rhs[col.off_s + i] : 0.0;
      vec[0][i] = v;
      if (i < ni) rscratch[col.off_s + i] = v;
    }
    // slot j & 1 holds r_j and c_{j-1} of layer j, read one layer ahead
    if (nz > 1) {
      RowAhead<SL, 2> ra;
      const double* src[2] = {rhs + col.off_s + ni, bsub};
      ra.load(src, ni);
      ra.store(ahead[1]);
    }
    __syncthreads();
    if constexpr (REFINE == 3) {
      if (nz > 1) load_tile<SL>(a, stk);
    }
    for (int j = 1; j < nz; ++j) {
      RowAhead<SL, 2> ra;
      if (j + 1 < nz) {
        const double* src[2] = {rhs + col.off_s + (int64_t)(j + 1) * ni, bsub + (int64_t)j * ni};
        ra.load(src, ni);
      }
      // REFINE 3: X_{j-1} was read during the previous layer; X_j is read
      // now, into an, and taken over once this layer is done
      double an[SL][SL];
      if constexpr (REFINE == 3) {
        if (j + 1 < nz) load_tile<SL>(an, stk + (int64_t)j * BLK);
      } else {
        load_tile<SL>(a, stk + (int64_t)(j - 1) * BLK);
      }
      tile_matvec<SL>(a, vec[(j - 1) & 1], w);
      if (j + 1 < nz) ra.store(ahead[(j + 1) & 1]);
      if (tc == 0) {
        const double* av = ahead[j & 1];
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int i = s * VK_G + tr;
          double r = 0.0;
          if (i < ni) {
            r = av[i] - av[NR + i] * w[s];
            rscratch[col.off_s + (int64_t)j * ni + i] = r;
          }
          vec[j & 1][i] = r;
        }
      }
      __syncthreads();
      if constexpr (REFINE == 3) {
        if (j + 1 < nz) {
#pragma unroll
          for (int s = 0; s < SL; ++s)
#pragma unroll
            for (int t = 0; t < SL; ++t) a[s][t] = an[s][t];
        }
      }
    }
    rm = rscratch;
    __syncthreads();
  }
  // Backward sweep.  vec holds t_j = r'_j - b_j o x_{j+1}.
  for (int i = tid; i < VK_G * SL; i += VK_NT)
    vec[(nz - 1) & 1][i] = (i < ni) ? rm[col.off_s + (int64_t)(nz - 1) * ni + i] : 0.0;
  __syncthreads();
  if constexpr (REFINE == 0) {
    for (int j = nz - 1; j >= 0; --j) {
      load_tile<SL>(a, stk + (int64_t)j * BLK);
      tile_matvec<SL>(a, vec[j & 1], w);
      if (tc == 0) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int i = s * VK_G + tr;
          if (i < ni) {
            const int64_t e = col.off_s + (int64_t)j * ni + i;
            const double xv = w[s];
            solve_epilogue(epi, e, xv, x);
            if (j > 0) {
              const int64_t e2 = col.off_s + (int64_t)(j - 1) * ni + i;
              vec[(j - 1) & 1][i] = rm[e2] - bsup[(int64_t)(j - 1) * ni + i] * xv;
            }
          } else if (j > 0) {
            vec[(j - 1) & 1][i] = 0.0;
          }
        }
      }
      __syncthreads();
    }
  } else if constexpr (REFINE == 2 || REFINE == 3) {
    // per layer: x0 = X_j t_j; r = t_j - A'_j x0; x_j = x0 + X_j r
    const double* sblk = dstk + col.off_b;
    bool wide = false;  // thread 0: some layer's correction exceeded its solution
    // slot j & 1: base, y, k1, yk2 of layer j, then r'_{j-1} and b_{j-1}
    auto ahead_src = [&](int j, const double* (&src)[6]) {
      const int64_t o = col.off_s + (int64_t)j * ni;
      src[0] = epi.base ? epi.base + o : nullptr;
      src[1] = epi.kind ? epi.y + o : nullptr;
      src[2] = epi.kind == 2 ? epi.k1 + o : nullptr;
      src[3] = epi.kind == 2 ? epi.yk2 + o : nullptr;
      src[4] = j > 0 ? rm + o - ni : nullptr;
      src[5] = j > 0 ? bsup + (int64_t)(j - 1) * ni : nullptr;
    };
    {
      RowAhead<SL, 6> ra;
      const double* src[6];
      ahead_src(nz - 1, src);
      ra.load(src, ni);
      ra.store(ahead[(nz - 1) & 1]);
    }
    __syncthreads();
    // REFINE 2: no block stays in registers across the layer; each of the
    // three products loads its whole block first, so all its loads are in
    // flight at once, and X_j is read again for the correction.
    // REFINE 3: X_j stays in registers for both of its products, and the
    // next layer's blocks are read while this one computes: the first
    // SLH tile rows of X_{j-1} at the start of the layer, A'_{j-1} once
    // A'_j is used, the other rows of X_{j-1} once X_j is used.
    constexpr int SLH = (SL + 1) / 2;
    double xh[SL][SL], sh2[SL][SL];
    if constexpr (REFINE == 3) {
      load_tile<SL>(xh, stk + (int64_t)(nz - 1) * BLK);
      load_tile<SL>(sh2, sblk + (int64_t)(nz - 1) * BLK);
    }
    for (int j = nz - 1; j >= 0; --j) {
      RowAhead<SL, 6> ra;
      if (j > 0) {
        const double* src[6];
        ahead_src(j - 1, src);
        ra.load(src, ni);
      }
      const double* xnext = stk + (int64_t)(j - 1) * BLK;  // X_{j-1} when j > 0
      double xn[SLH][SL];
      if constexpr (REFINE == 3) {
        if (j > 0) {
#pragma unroll
          for (int s = 0; s < SLH; ++s)
#pragma unroll
            for (int t = 0; t < SL; ++t) xn[s][t] = xnext[(s * SL + t) * VK_NT + tid];
        }
        tile_matvec<SL>(xh, vec[j & 1], w);                   // x0_j
      } else {
        double t0[SL][SL];
        load_tile<SL>(t0, stk + (int64_t)j * BLK);
        tile_matvec<SL>(t0, vec[j & 1], w);                   // x0_j
      }
      if (j > 0) ra.store(ahead[(j - 1) & 1]);
      if (tc == 0) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int i = s * VK_G + tr;
          rv[i] = (i < ni) ? w[s] : 0.0;
        }
      }
      __syncthreads();
      double q[SL];
      if constexpr (REFINE == 3) {
        tile_matvec<SL>(sh2, rv, q);                          // A'_j x0_j
        if (j > 0) load_tile<SL>(sh2, sblk + (int64_t)(j - 1) * BLK);
      } else {
        double t1[SL][SL];
        load_tile<SL>(t1, sblk + (int64_t)j * BLK);
        tile_matvec<SL>(t1, rv, q);                           // A'_j x0_j
      }
      if (tc == 0) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int i = s * VK_G + tr;
          bx[i] = (i < ni) ? vec[j & 1][i] - q[s] : 0.0;
        }
      }
      __syncthreads();
      double d[SL];
      if constexpr (REFINE == 3) {
        tile_matvec<SL>(xh, bx, d);                           // X_j (t_j - A'_j x0_j)
        if (j > 0) {
#pragma unroll
          for (int s = 0; s < SL; ++s)
#pragma unroll
            for (int t = 0; t < SL; ++t)
              xh[s][t] = s < SLH ? xn[s < SLH ? s : 0][t] : xnext[(s * SL + t) * VK_NT + tid];
        }
      } else {
        // an opaque copy of the address, so the compiler reloads X_j
        // instead of holding the first load's registers across the layer
        const double* xj = stk + (int64_t)j * BLK;
        asm volatile("" : "+s"(xj));
        double t2[SL][SL];
        load_tile<SL>(t2, xj);
        tile_matvec<SL>(t2, bx, d);                           // X_j (t_j - A'_j x0_j)
      }
      double dmax = 0.0, xmax = 0.0;
      if (tc == 0) {
        const double* av = ahead[j & 1];
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int i = s * VK_G + tr;
          if (i < ni) {
            const double xv = rv[i] + d[s];  // rv still holds x0_j (w's registers are free)
            dmax = fmax(dmax, fabs(d[s]));
            xmax = fmax(xmax, fabs(xv));
            solve_epilogue_v(epi, col.off_s + (int64_t)j * ni + i, xv, x, av[i], av[NR + i],
                             av[2 * NR + i], av[3 * NR + i]);
            if (j > 0) vec[(j - 1) & 1][i] = av[4 * NR + i] - av[5 * NR + i] * xv;
          } else if (j > 0) {
            vec[(j - 1) & 1][i] = 0.0;
          }
        }
      }
      if (refl) {
        // the tc == 0 lanes of a wave are 0, 16, 32, 48
        dmax = fmax(dmax, __shfl_xor(dmax, 16));
        dmax = fmax(dmax, __shfl_xor(dmax, 32));
        xmax = fmax(xmax, __shfl_xor(xmax, 16));
        xmax = fmax(xmax, __shfl_xor(xmax, 32));
        if ((tid & 63) == 0) {
          red[tid >> 6][0] = dmax;
          red[tid >> 6][1] = xmax;
        }
      }
      __syncthreads();
      if (refl && tid == 0) {
        double dm = red[0][0], xm = red[0][1];
#pragma unroll
        for (int v = 1; v < VK_NT / 64; ++v) {
          dm = fmax(dm, red[v][0]);
          xm = fmax(xm, red[v][1]);
        }
        wide = wide || dm > xm;
      }
    }
    if (refl && tid == 0) refl[c] = wide ? 0 : VK_ST_DONE;
  } else {
    // step j: load X_j; finish layer u = j + 1 (its residual needs X_j, its
    // correction X_u, kept in xp); form x0_j = X_j t_j.  rv holds x0_u, bx
    // holds b_j o x0_u, w (registers) x0_u.
    const double* dblk = dstk + col.off_b;
    double xp[SL][SL];
    bool wide = false;  // thread 0: some layer's correction exceeded its solution
    for (int j = nz - 1; j >= -1; --j) {
      if (j >= 0) load_tile<SL>(a, stk + (int64_t)j * BLK);
      if (j < nz - 1) {
        const int u = j + 1;
        double q[SL], pz[SL];
        stream_matvec<SL>(dblk + (int64_t)u * BLK, rv, q);   // D_u x0_u
        if (j >= 0) tile_matvec<SL>(a, bx, pz);              // X_j (b_j o x0_u)
        __syncthreads();
        if (tc == 0) {
#pragma unroll
          for (int s = 0; s < SL; ++s) {
            const int i = s * VK_G + tr;
            double r = 0.0;
            if (i < ni) {
              double ax = q[s];
              if (j >= 0) ax = ax - bsub[(int64_t)j * ni + i] * pz[s];
              r = vec[u & 1][i] - ax;
            }
            rv[i] = r;
          }
        }
        __syncthreads();
        double d[SL];
        tile_matvec<SL>(xp, rv, d);                          // X_u (t_u - A'_u x0_u)
        double dmax = 0.0, xmax = 0.0;
        if (tc == 0) {
#pragma unroll
          for (int s = 0; s < SL; ++s) {
            const int i = s * VK_G + tr;
            if (i < ni) {
              const double xv = w[s] + d[s];
              dmax = fmax(dmax, fabs(d[s]));
              xmax = fmax(xmax, fabs(xv));
              solve_epilogue(epi, col.off_s + (int64_t)u * ni + i, xv, x);
              if (j >= 0)
                vec[j & 1][i] = rm[col.off_s + (int64_t)j * ni + i] -
                                bsup[(int64_t)j * ni + i] * xv;
            } else if (j >= 0) {
              vec[j & 1][i] = 0.0;
            }
          }
        }
        if (refl) {
          // the tc == 0 lanes of a wave are 0, 16, 32, 48
          dmax = fmax(dmax, __shfl_xor(dmax, 16));
          dmax = fmax(dmax, __shfl_xor(dmax, 32));
          xmax = fmax(xmax, __shfl_xor(xmax, 16));
          xmax = fmax(xmax, __shfl_xor(xmax, 32));
          if ((tid & 63) == 0) {
            red[tid >> 6][0] = dmax;
            red[tid >> 6][1] = xmax;
          }
        }
        __syncthreads();
        if (refl && tid == 0) {
          double dm = red[0][0], xm = red[0][1];
#pragma unroll
          for (int v = 1; v < VK_NT / 64; ++v) {
            dm = fmax(dm, red[v][0]);
            xm = fmax(xm, red[v][1]);
          }
          wide = wide || dm > xm;
        }
      }
      if (j < 0) break;
      tile_matvec<SL>(a, vec[j & 1], w);                     // x0_j
      if (tc == 0) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int i = s * VK_G + tr;
          rv[i] = (i < ni) ? w[s] : 0.0;
          bx[i] = (j > 0 && i < ni) ? bsup[(int64_t)(j - 1) * ni + i] * w[s] : 0.0;
        }
      }
      __syncthreads();
#pragma unroll
      for (int s = 0; s < SL; ++s)
#pragma unroll
        for (int t = 0; t < SL; ++t) xp[s][t] = a[s][t];
    }
    if (refl && tid == 0) refl[c] = wide ? 0 : VK_ST_DONE;
  }
}

// Residual of the block-tridiagonal system at x, one workgroup per (layer,
// column):  res_j = ((r_j - D_j x_j) - b_j o x_{j+1}) - c_{j-1} o x_{j-1},
// D_j the tile-layout diagonal blocks, b = sup, c = sub.  The refinement
// step of the cyclic-reduction solves.
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_bt_residual(
    const VkCol* __restrict__ cols, int ncols, const double* __restrict__ blk,
    const double* __restrict__ sup_d, const double* __restrict__ sub_d,
    const double* __restrict__ rhs, const double* __restrict__ x, double* __restrict__ res,
    const int* __restrict__ active) {
  __shared__ double xv[VK_G * SL];
  const int c = blockIdx.y, j = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL || j >= col.nz) return;
  const int nz = col.nz, ni = col.ni, tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const int64_t row = col.off_s + (int64_t)j * ni;
  for (int i = tid; i < VK_G * SL; i += VK_NT) xv[i] = (i < ni) ? x[row + i] : 0.0;
  __syncthreads();
  double q[SL];
  stream_matvec<SL>(blk + col.off_b + (int64_t)j * vk_blk<SL>(), xv, q);
  if (tc == 0) {
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int i = s * VK_G + tr;
      if (i < ni) {
        double r = rhs[row + i] - q[s];
        if (j + 1 < nz) r = r - sup_d[row + i] * x[row + ni + i];
        if (j > 0) r = r - sub_d[row - ni + i] * x[row - ni + i];
        res[row + i] = r;
      }
    }
  }
}

// Dense block-tridiagonal Thomas (solver.py:129-176): full ni x ni
// off-diagonal blocks, one workgroup per column, one right-hand side.
//   A'_0 = D_0;  A'_j = D_j - C_j X_{j-1} B_{j-1};  X_j = inv(A'_j)
//   r'_0 = r_0;  r'_j = r_j - C_j (X_{j-1} r'_{j-1})
//   x_{nz-1} = X_{nz-1} r'_{nz-1};  x_j = X_j (r'_j - B_j x_{j+1})
// with B_j = sup[j], C_j = sub[j-1] as in the reference's scan.  diag, sup
// and sub are natural (ni, ni) blocks at off_s * ni, layer j at j ni^2 (sup
// and sub padded to nz blocks per column, the last unused); rhs, x and rmod
// (r'_j) at off_s; inv receives X_j in tile layout at off_b.  The product
// C_j X_{j-1} B_{j-1} is formed as (C_j X_{j-1}) B_{j-1}, X_{j-1} and then
// C_j X_{j-1} staged in LDS.
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_bt_dense(
    const VkCol* __restrict__ cols, int ncols, const double* __restrict__ diag,
    const double* __restrict__ sup, const double* __restrict__ sub,
    const double* __restrict__ rhs, double* __restrict__ x, double* __restrict__ inv,
    double* __restrict__ rmod, int* __restrict__ status) {
  constexpr int N = VK_G * SL;
  __shared__ BtShared<SL> sh;
  __shared__ double xs[N * N];   // X_{j-1}, then C_j X_{j-1} (row-major, padded)
  __shared__ double wv[N];       // X_{j-1} r'_{j-1}, then the backward sweep's vector
  const int c = blockIdx.x;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const int nz = col.nz, ni = col.ni;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  constexpr int BLK = vk_blk<SL>();
  const int64_t nn = (int64_t)ni * ni;
  const double* D = diag + col.off_s * ni;
  const double* Bs = sup + col.off_s * ni;
  const double* Cs = sub + col.off_s * ni;
  double a[SL][SL];
  int bad = 0;
  for (int j = 0; j < nz; ++j) {
    const double* Dj = D + j * nn;
#pragma unroll
    for (int s = 0; s < SL; ++s)
#pragma unroll
      for (int t = 0; t < SL; ++t) a[s][t] = nat_elem(Dj, ni, s * VK_G + tr, t * VK_G + tc, 0);
    if (j == 0) {
      for (int i = tid; i < N; i += VK_NT) {
        const double v = (i < ni) ? rhs[col.off_s + i] : 0.0;
        sh.vec[0][i] = v;
        if (i < ni) rmod[col.off_s + i] = v;
      }
    } else {
      const double* Cj = Cs + (int64_t)(j - 1) * nn;   // C_j = sub[j-1]
      const double* Bj = Bs + (int64_t)(j - 1) * nn;   // B_{j-1} = sup[j-1]
      // r'_j = r_j - C_j w, w = X_{j-1} r'_{j-1} in wv
      double cw[SL];
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int i = s * VK_G + tr;
        double acc = 0.0;
#pragma unroll
        for (int t = 0; t < SL; ++t) acc = fma(nat_elem(Cj, ni, i, t * VK_G + tc, 0),
                                               wv[t * VK_G + tc], acc);
        cw[s] = vk_rowsum16(acc);
      }
      if (tc == 0) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int i = s * VK_G + tr;
          double r = 0.0;
          if (i < ni) {
            r = rhs[col.off_s + (int64_t)j * ni + i] - cw[s];
            rmod[col.off_s + (int64_t)j * ni + i] = r;
          }
          sh.vec[j & 1][i] = r;
        }
      }
      // Q = C_j X_{j-1} (X_{j-1} in xs)
      double q[SL][SL];
#pragma unroll
      for (int s = 0; s < SL; ++s)
#pragma unroll
        for (int t = 0; t < SL; ++t) q[s][t] = 0.0;
      for (int k = 0; k < ni; ++k) {
        double ck[SL], xk[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) ck[s] = nat_elem(Cj, ni, s * VK_G + tr, k, 0);
#pragma unroll
        for (int t = 0; t < SL; ++t) xk[t] = xs[k * N + t * VK_G + tc];
#pragma unroll
        for (int s = 0; s < SL; ++s)
#pragma unroll
          for (int t = 0; t < SL; ++t) q[s][t] = fma(ck[s], xk[t], q[s][t]);
      }
      __syncthreads();  // every thread is done with X_{j-1} in xs
#pragma unroll
      for (int s = 0; s < SL; ++s)
#pragma unroll
        for (int t = 0; t < SL; ++t) xs[(s * VK_G + tr) * N + t * VK_G + tc] = q[s][t];
      __syncthreads();
      // A'_j = D_j - Q B_{j-1}
      for (int k = 0; k < ni; ++k) {
        double qk[SL], bk[SL];
#pragma unroll
        for (int s = 0; s < SL; ++s) qk[s] = xs[(s * VK_G + tr) * N + k];
#pragma unroll
        for (int t = 0; t < SL; ++t) bk[t] = nat_elem(Bj, ni, k, t * VK_G + tc, 0);
#pragma unroll
        for (int s = 0; s < SL; ++s)
#pragma unroll
          for (int t = 0; t < SL; ++t) a[s][t] = fma(-qk[s], bk[t], a[s][t]);
      }
    }
    __syncthreads();
    bad |= gj_invert<SL>(a, ni, sh);
    store_tile<SL>(a, inv + col.off_b + (int64_t)j * BLK);
    // X_j into xs for the next layer; w = X_j r'_j into wv
    double w[SL];
    tile_matvec<SL>(a, sh.vec[j & 1], w);
#pragma unroll
    for (int s = 0; s < SL; ++s)
#pragma unroll
      for (int t = 0; t < SL; ++t) xs[(s * VK_G + tr) * N + t * VK_G + tc] = a[s][t];
    if (tc == 0) {
#pragma unroll
      for (int s = 0; s < SL; ++s) wv[s * VK_G + tr] = (s * VK_G + tr < ni) ? w[s] : 0.0;
    }
    __syncthreads();
  }
  // backward sweep: x_{nz-1} = X_{nz-1} r'_{nz-1} (= wv now)
  for (int i = tid; i < ni; i += VK_NT) x[col.off_s + (int64_t)(nz - 1) * ni + i] = wv[i];
  __syncthreads();
  for (int j = nz - 2; j >= 0; --j) {
    // t_j = r'_j - B_j x_{j+1} with x_{j+1} in wv, into sh.vec[0]
    const double* Bj = Bs + (int64_t)j * nn;
    double bx[SL];
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int i = s * VK_G + tr;
      double acc = 0.0;
#pragma unroll
      for (int t = 0; t < SL; ++t) acc = fma(nat_elem(Bj, ni, i, t * VK_G + tc, 0),
                                             wv[t * VK_G + tc], acc);
      bx[s] = vk_rowsum16(acc);
    }
    if (tc == 0) {
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int i = s * VK_G + tr;
        sh.vec[0][i] = (i < ni) ? rmod[col.off_s + (int64_t)j * ni + i] - bx[s] : 0.0;
      }
    }
    __syncthreads();
    double xv[SL];
    stream_matvec<SL>(inv + col.off_b + (int64_t)j * BLK, sh.vec[0], xv);
    __syncthreads();  // every thread has read wv and sh.vec[0]
    if (tc == 0) {
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int i = s * VK_G + tr;
        if (i < ni) {
          x[col.off_s + (int64_t)j * ni + i] = xv[s];
          wv[i] = xv[s];
        } else {
          wv[i] = 0.0;
        }
      }
    }
    __syncthreads();
  }
  if (bad && tid == 0 && status) atomicOr(&status[c], VK_ST_SINGULAR);
}

// Copy a tile-layout inverse stack into natural (nz, ni, ni) blocks.
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_tile_to_nat(const VkCol* __restrict__ cols, int ncols,
                                                       const double* __restrict__ tile,
                                                       double* __restrict__ nat) {
  const int c = blockIdx.y;
  const int j = blockIdx.x;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL || j >= col.nz) return;
  const int ni = col.ni, tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const double* src = tile + col.off_b + (int64_t)j * vk_blk<SL>();
  double* dst = nat + col.off_s * ni + (int64_t)j * ni * ni;
  for (int s = 0; s < SL; ++s)
    for (int t = 0; t < SL; ++t) {
      const int i = s * VK_G + tr, q = t * VK_G + tc;
      if (i < ni && q < ni) dst[(int64_t)i * ni + q] = src[(s * SL + t) * VK_NT + tid];
    }
}

// ---------------------------------------------------------------------------
// Block cyclic reduction: the same block-tridiagonal system, for batches of a
// few columns, where one workgroup per column leaves the GPU idle.  Level l
// holds n_l block rows (n_0 = nz, n_{l+1} = ceil(n_l / 2)).  Each odd row j
// is eliminated (its diagonal block inverted with the register Gauss-Jordan
// and its partial pivoting), giving P_j = U_{j-1} inv(D_j) and
// Q_j = L_{j+1} inv(D_j); each even row i becomes a row of level l+1:
//   D' = (D_i - Q_{i-1} U_{i-1}) - P_{i+1} L_{i+1},
//   L' = -Q_{i-1} L_{i-1},  U' = -P_{i+1} U_{i+1}.
// The solve reduces r the same way, solves the single top row and
// substitutes back level by level:  x_j = inv(D_j) (r_j - L_j x_{j-1} -
// U_j x_{j+1}) for the odd rows.  Every level is one launch whose
// workgroups each own one block row, so the critical path is about
// log2(nz) block inversions instead of nz.  Level 0 reads the diagonal
// blocks from the tile-layout stack the chemistry writes and keeps L, U as
// the diagonal vectors sub_d / sup_d; later levels store dense blocks, all
// in the register-tile layout, in a region of 7 nz N^2 doubles per column
// at 7 * off_b (N = 16 SL; the levels need at most 6.43 nz N^2).
// ---------------------------------------------------------------------------

namespace {

struct CrLvl {
  int n, h;                           // block rows, eliminated (odd) rows
  int64_t D, L, U, Dinv, P, Q, r;     // offsets (doubles) in the region
};

// Offsets of level `level` of a column with nz rows; past the top level the
// top level is returned (n == 1).
__device__ inline CrLvl cr_level(int nz, int64_t blk, int N, int64_t base, int level) {
  CrLvl v{};
  int n = nz;
  int64_t off = base;
  for (int l = 0;; ++l) {
    v.n = n;
    v.h = n / 2;
    v.D = v.L = v.U = v.P = v.Q = -1;
    if (l >= 1) {
      v.D = off;
      off += (int64_t)n * blk;
      if (n > 1) {
        v.L = off;
        off += (int64_t)n * blk;
        v.U = off;
        off += (int64_t)n * blk;
      }
    }
    v.Dinv = off;
    if (n > 1) {
      off += (int64_t)v.h * blk;
      v.P = off;
      off += (int64_t)v.h * blk;
      v.Q = off;
      off += (int64_t)v.h * blk;
    } else {
      off += blk;
    }
    v.r = off;
    off += (int64_t)n * N;
    if (l == level || n == 1) return v;
    n = (n + 1) / 2;
  }
}

__device__ inline int cr_top_level(int nz) {
  int l = 0;
  for (int n = nz; n > 1; n = (n + 1) / 2) ++l;
  return l;
}

// Diagonal block of row j at `level` into the register tile, zero outside ni.
template <int SL>
__device__ inline void cr_load_diag(double (&a)[SL][SL], const VkCol& col, const double* blk0,
                                    const double* cr, const CrLvl& v, int level, int j) {
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const double* src = level == 0 ? blk0 + col.off_b + (int64_t)j * vk_blk<SL>()
                                 : cr + v.D + (int64_t)j * vk_blk<SL>();
  load_tile<SL>(a, src);
#pragma unroll
  for (int s = 0; s < SL; ++s)
#pragma unroll
    for (int t = 0; t < SL; ++t)
      if (s * VK_G + tr >= col.ni || t * VK_G + tc >= col.ni) a[s][t] = 0.0;
}

// Dense products on FP64 MFMA (v_mfma_f64_16x16x4f64).  Output tile t =
// (a, b) of the N x N result belongs to wave t mod 4, as T[t / 4] in the
// MFMA result layout: lane l, item v holds element (16a + l/16 + 4v,
// 16b + l%16).  Operands are tile-layout blocks in global memory, staged in
// 16-wide panels: ga[row][17] (padded against bank conflicts), gb[k][N].
typedef double cr_d4 __attribute__((ext_vector_type(4)));
template <int SL>
constexpr int cr_tpw() { return (SL * SL + 3) / 4; }

template <int SL>
__device__ inline void cr_mm_zero(cr_d4 (&T)[cr_tpw<SL>()]) {
#pragma unroll
  for (int s = 0; s < cr_tpw<SL>(); ++s) T[s] = cr_d4{0.0, 0.0, 0.0, 0.0};
}

template <int SL>
__device__ inline void cr_mm_load(cr_d4 (&T)[cr_tpw<SL>()], const double* __restrict__ src) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane >> 4, lc = lane & 15;
#pragma unroll
  for (int s = 0; s < cr_tpw<SL>(); ++s) {
    const int t = wave + 4 * s;
    if (t < SL * SL) {
#pragma unroll
      for (int v = 0; v < 4; ++v) T[s][v] = src[t * VK_NT + (lr + 4 * v) * VK_G + lc];
    }
  }
}

template <int SL>
__device__ inline void cr_mm_store(const cr_d4 (&T)[cr_tpw<SL>()], double* __restrict__ dst) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane >> 4, lc = lane & 15;
#pragma unroll
  for (int s = 0; s < cr_tpw<SL>(); ++s) {
    const int t = wave + 4 * s;
    if (t < SL * SL) {
#pragma unroll
      for (int v = 0; v < 4; ++v) dst[t * VK_NT + (lr + 4 * v) * VK_G + lc] = T[s][v];
    }
  }
}

// One 16-wide panel of T += sgn * A B from ga / gb.
template <int SL>
__device__ inline void cr_mm_panel(double sgn, cr_d4 (&T)[cr_tpw<SL>()], const double* ga,
                                   const double* gb) {
  constexpr int N = VK_G * SL, GA = VK_G + 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane >> 4, lc = lane & 15;
#pragma unroll
  for (int s = 0; s < cr_tpw<SL>(); ++s) {
    const int t = wave + 4 * s;
    if (t < SL * SL) {
      const int a = t / SL, b = t - a * SL;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double av = sgn * ga[(a * VK_G + lc) * GA + 4 * q + lr];
        const double bw = gb[(4 * q + lr) * N + b * VK_G + lc];
        T[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bw, T[s], 0, 0, 0);
      }
    }
  }
}

// T += sgn * A B (sgn = +1 or -1, exact).  PRE: both operands are loaded
// into registers first, so their latencies overlap (a workgroup alone on its
// CU has nothing else to hide them behind); it costs 144 registers, which
// the update kernel has to spare and the elimination kernel (two waves per
// SIMD) does not, so there the panels are loaded one at a time.
template <int SL, bool PRE>
__device__ inline void cr_mm(const double* __restrict__ A, const double* __restrict__ B,
                             double sgn, cr_d4 (&T)[cr_tpw<SL>()], double* ga, double* gb) {
  constexpr int N = VK_G * SL, GA = VK_G + 1;
  const int tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  if constexpr (PRE) {
    double ra[SL][SL], rb[SL][SL];
    load_tile<SL>(ra, A);
    load_tile<SL>(rb, B);
#pragma unroll
    for (int p = 0; p < SL; ++p) {
#pragma unroll
      for (int s = 0; s < SL; ++s) ga[(s * VK_G + tr) * GA + tc] = ra[s][p];
#pragma unroll
      for (int u = 0; u < SL; ++u) gb[tr * N + u * VK_G + tc] = rb[p][u];
      __syncthreads();
      cr_mm_panel<SL>(sgn, T, ga, gb);
      __syncthreads();
    }
  } else {
#pragma unroll 1
    for (int p = 0; p < SL; ++p) {
#pragma unroll
      for (int s = 0; s < SL; ++s) ga[(s * VK_G + tr) * GA + tc] = A[(s * SL + p) * VK_NT + tid];
#pragma unroll
      for (int u = 0; u < SL; ++u) gb[tr * N + u * VK_G + tc] = B[(p * SL + u) * VK_NT + tid];
      __syncthreads();
      cr_mm_panel<SL>(sgn, T, ga, gb);
      __syncthreads();
    }
  }
}

__device__ inline void cr_epilogue(const SolveEpi& epi, int64_t e, double xv, double* x) {
  if (epi.base) xv = epi.base[e] + xv;
  x[e] = xv;
  if (epi.kind == 1) {
    epi.out1[e] = epi.y[e] + xv / epi.e0;
  } else if (epi.kind == 2) {
#pragma clang fp contract(off)
    const double sol = epi.y[e] + epi.e0 * epi.k1[e] + epi.e1 * xv;
    epi.out1[e] = sol;
    epi.out2[e] = fabs(sol - epi.yk2[e]);
  }
}

}  // namespace

// Eliminate the odd rows of one level: inv(D_j), P_j, Q_j.  Grid (h, ncols).
// WIDE: compiled for four workgroups per CU (128 VGPRs; what spills lies
// outside the pivot steps), for levels with many blocks to invert; else
// unconstrained, for the few-block levels whose time is one inversion's
// latency.
template <int SL, bool WIDE>
__global__ __launch_bounds__(VK_NT) __attribute__((amdgpu_waves_per_eu(WIDE && SL <= 6 ? 4 : 1))) void k_cr_elim(const VkCol* __restrict__ cols, int ncols,
                                                    const double* __restrict__ blk0,
                                                    const double* __restrict__ sup_d,
                                                    const double* __restrict__ sub_d,
                                                    double* __restrict__ cr, int level,
                                                    int* __restrict__ status,
                                                    const int* __restrict__ active) {
  // The inversion's shared state and the product panels are never live
  // together (a barrier parts them), so they share their LDS: four
  // workgroups fit a CU, as the registers allow, instead of three.
  __shared__ union {
    BtShared<SL> sh;
    struct {
      double a[(VK_G + 1) * VK_G * SL], b[VK_G * VK_G * SL];
    } g;
  } lds;
  BtShared<SL>& sh = lds.sh;
  double* ga = lds.g.a;
  double* gb = lds.g.b;
  const int c = blockIdx.y, w = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  constexpr int BLK = vk_blk<SL>(), N = VK_G * SL;
  const CrLvl v = cr_level(col.nz, BLK, N, 7 * col.off_b, level);
  if (v.n <= 1 || w >= v.h) return;
  const int ni = col.ni, tid = threadIdx.x, tr = tid >> 4;
  const int j = 2 * w + 1;
  double a[SL][SL];
  cr_load_diag<SL>(a, col, blk0, cr, v, level, j);
  const int bad = gj_invert<SL>(a, ni, sh);
  if (bad && tid == 0 && status) atomicOr(&status[c], VK_ST_SINGULAR);
  store_tile<SL>(a, cr + v.Dinv + (int64_t)w * BLK);
  const bool has_q = j + 1 < v.n;
  if (level == 0) {
    // U_{j-1} = diag(sup_d[j-1]), L_{j+1} = diag(sub_d[j])
    const double* su = sup_d + col.off_s + (int64_t)(j - 1) * ni;
    const double* sl = sub_d + col.off_s + (int64_t)j * ni;
    double p[SL][SL];
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int i = s * VK_G + tr;
      const double u = i < ni ? su[i] : 0.0;
#pragma unroll
      for (int t = 0; t < SL; ++t) p[s][t] = u * a[s][t];
    }
    store_tile<SL>(p, cr + v.P + (int64_t)w * BLK);
    if (has_q) {
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        const int i = s * VK_G + tr;
        const double lv = i < ni ? sl[i] : 0.0;
#pragma unroll
        for (int t = 0; t < SL; ++t) p[s][t] = lv * a[s][t];
      }
      store_tile<SL>(p, cr + v.Q + (int64_t)w * BLK);
    }
  } else {
    __syncthreads();  // inv(D_j) is read back through the panels
    const double* Dinv = cr + v.Dinv + (int64_t)w * BLK;
    cr_d4 T[cr_tpw<SL>()];
    cr_mm_zero<SL>(T);
    cr_mm<SL, false>(cr + v.U + (int64_t)(j - 1) * BLK, Dinv, 1.0, T, ga, gb);
    cr_mm_store<SL>(T, cr + v.P + (int64_t)w * BLK);
    if (has_q) {
      cr_mm_zero<SL>(T);
      cr_mm<SL, false>(cr + v.L + (int64_t)(j + 1) * BLK, Dinv, 1.0, T, ga, gb);
      cr_mm_store<SL>(T, cr + v.Q + (int64_t)w * BLK);
    }
  }
}

// Even rows of one level become the rows of the next.  Grid (ceil(n/2), ncols).
// Level 0 scales by the diffusion diagonals; later levels multiply dense
// blocks on MFMA.
// WIDE (levels with many rows, as k_cr_elim): the products load their
// operands panel by panel through LDS instead of both whole blocks into
// registers first, which leaves room for two workgroups per CU.
template <int SL, bool WIDE>
__global__ __launch_bounds__(VK_NT) __attribute__((amdgpu_waves_per_eu(WIDE && SL <= 6 ? 2 : 1))) void k_cr_update(const VkCol* __restrict__ cols, int ncols,
                                                      const double* __restrict__ blk0,
                                                      const double* __restrict__ sup_d,
                                                      const double* __restrict__ sub_d,
                                                      double* __restrict__ cr, int level,
                                                      const double* __restrict__ rhs,
                                                      const int* __restrict__ active) {
  __shared__ double ga[(VK_G + 1) * VK_G * SL], gb[VK_G * VK_G * SL];
  __shared__ double rv[3][VK_G * SL];
  const int c = blockIdx.y, w = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  constexpr int BLK = vk_blk<SL>(), N = VK_G * SL;
  const CrLvl v = cr_level(col.nz, BLK, N, 7 * col.off_b, level);
  const int i = 2 * w;
  if (v.n <= 1 || i >= v.n) return;
  const CrLvl vn = cr_level(col.nz, BLK, N, 7 * col.off_b, level + 1);
  // the stage-1 right-hand side's reduction, which needs only this level's
  // P and Q (written by k_cr_elim before this launch)
  if (rhs) cr_fwd_row<SL>(col, v, vn, level, w, rhs, cr, rv);
  const bool has_p = i + 1 < v.n, has_u = i + 2 < v.n;
  if (level >= 1) {
    // D' = (D_i - Q_{i-1} U_{i-1}) - P_{i+1} L_{i+1}; L' = -Q_{i-1} L_{i-1};
    // U' = -P_{i+1} U_{i+1}
    cr_d4 T[cr_tpw<SL>()], X[cr_tpw<SL>()];
    cr_mm_load<SL>(T, cr + v.D + (int64_t)i * BLK);
    if (i >= 1) {
      const double* Qp = cr + v.Q + (int64_t)(w - 1) * BLK;
      cr_mm<SL, !WIDE>(Qp, cr + v.U + (int64_t)(i - 1) * BLK, -1.0, T, ga, gb);
      if (i >= 2) {
        cr_mm_zero<SL>(X);
        cr_mm<SL, !WIDE>(Qp, cr + v.L + (int64_t)(i - 1) * BLK, -1.0, X, ga, gb);
        cr_mm_store<SL>(X, cr + vn.L + (int64_t)w * BLK);
      }
    }
    if (has_p) {
      const double* Pp = cr + v.P + (int64_t)w * BLK;
      cr_mm<SL, !WIDE>(Pp, cr + v.L + (int64_t)(i + 1) * BLK, -1.0, T, ga, gb);
      if (has_u) {
        cr_mm_zero<SL>(X);
        cr_mm<SL, !WIDE>(Pp, cr + v.U + (int64_t)(i + 1) * BLK, -1.0, X, ga, gb);
        cr_mm_store<SL>(X, cr + vn.U + (int64_t)w * BLK);
      }
    }
    cr_mm_store<SL>(T, cr + vn.D + (int64_t)w * BLK);
    return;
  }
  const int ni = col.ni, tid = threadIdx.x, tc = tid & 15;
  double d[SL][SL], m[SL][SL];
  cr_load_diag<SL>(d, col, blk0, cr, v, level, i);
  if (i >= 1) {
    // odd row i-1: Q_{i-1} = L_i inv(D_{i-1}); U_{i-1} = diag(sup_d[i-1]),
    // L_{i-1} = diag(sub_d[i-2]); m becomes L'
    load_tile<SL>(m, cr + v.Q + (int64_t)(w - 1) * BLK);
    const double* bu = sup_d + col.off_s + (int64_t)(i - 1) * ni;
    const double* bl = sub_d + col.off_s + (int64_t)(i >= 2 ? i - 2 : 0) * ni;
#pragma unroll
    for (int t = 0; t < SL; ++t) {
      const int q = t * VK_G + tc;
      const double b = q < ni ? bu[q] : 0.0;
      const double b2 = (q < ni && i >= 2) ? bl[q] : 0.0;
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        d[s][t] = d[s][t] - m[s][t] * b;
        m[s][t] = -(m[s][t] * b2);
      }
    }
    if (i >= 2) store_tile<SL>(m, cr + vn.L + (int64_t)w * BLK);
  }
  if (has_p) {
    // odd row i+1: P_{i+1} = U_i inv(D_{i+1}); L_{i+1} = diag(sub_d[i]),
    // U_{i+1} = diag(sup_d[i+1]); m becomes U'
    load_tile<SL>(m, cr + v.P + (int64_t)w * BLK);
    const double* bl = sub_d + col.off_s + (int64_t)i * ni;
    const double* bu = sup_d + col.off_s + (int64_t)(has_u ? i + 1 : i) * ni;
#pragma unroll
    for (int t = 0; t < SL; ++t) {
      const int q = t * VK_G + tc;
      const double b = q < ni ? bl[q] : 0.0;
      const double b2 = (q < ni && has_u) ? bu[q] : 0.0;
#pragma unroll
      for (int s = 0; s < SL; ++s) {
        d[s][t] = d[s][t] - m[s][t] * b;
        m[s][t] = -(m[s][t] * b2);
      }
    }
    if (has_u) store_tile<SL>(m, cr + vn.U + (int64_t)w * BLK);
  }
  store_tile<SL>(d, cr + vn.D + (int64_t)w * BLK);
}

// Invert the single row of the top level.  Grid (ncols).
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_cr_top(const VkCol* __restrict__ cols, int ncols,
                                                   const double* __restrict__ blk0,
                                                   double* __restrict__ cr,
                                                   int* __restrict__ status,
                                                   const int* __restrict__ active) {
  __shared__ BtShared<SL> sh;
  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  constexpr int BLK = vk_blk<SL>(), N = VK_G * SL;
  const int top = cr_top_level(col.nz);
  const CrLvl v = cr_level(col.nz, BLK, N, 7 * col.off_b, top);
  double a[SL][SL];
  cr_load_diag<SL>(a, col, blk0, cr, v, top, 0);
  const int bad = gj_invert<SL>(a, col.ni, sh);
  if (bad && threadIdx.x == 0 && status) atomicOr(&status[c], VK_ST_SINGULAR);
  store_tile<SL>(a, cr + v.Dinv);
}

namespace {

// r_{l+1}[w] = (r_i - Q_{i-1} r_{i-1}) - P_{i+1} r_{i+1} for even row i = 2w
// of level `level` (level 0 reads r from rhs).  Shared by k_cr_fwd and, for
// the stage-1 right-hand side, by k_cr_update, so both form it alike.
template <int SL>
__device__ inline void cr_fwd_row(const VkCol& col, const CrLvl& v, const CrLvl& vn, int level,
                                  int w, const double* __restrict__ rhs, double* __restrict__ cr,
                                  double (&rv)[3][VK_G * SL]) {
  constexpr int BLK = vk_blk<SL>(), N = VK_G * SL;
  const int i = 2 * w;
  const int ni = col.ni, tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  for (int e = tid; e < 3 * N; e += VK_NT) {
    const int row = i - 1 + e / N, q = e % N;
    double x = 0.0;
    if (row >= 0 && row < v.n && q < ni)
      x = level == 0 ? rhs[col.off_s + (int64_t)row * ni + q] : cr[v.r + (int64_t)row * N + q];
    rv[e / N][q] = x;
  }
  __syncthreads();
  double out[SL], y[SL], a[SL][SL];
#pragma unroll
  for (int s = 0; s < SL; ++s) out[s] = rv[1][s * VK_G + tr];
  if (i >= 1) {
    load_tile<SL>(a, cr + v.Q + (int64_t)(w - 1) * BLK);
    tile_matvec<SL>(a, rv[0], y);
#pragma unroll
    for (int s = 0; s < SL; ++s) out[s] = out[s] - y[s];
  }
  if (i + 1 < v.n) {
    load_tile<SL>(a, cr + v.P + (int64_t)w * BLK);
    tile_matvec<SL>(a, rv[2], y);
#pragma unroll
    for (int s = 0; s < SL; ++s) out[s] = out[s] - y[s];
  }
  if (tc == 0) {
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int q = s * VK_G + tr;
      cr[vn.r + (int64_t)w * N + q] = q < ni ? out[s] : 0.0;
    }
  }
}

}  // namespace

// Reduce r: r'_i = (r_i - Q_{i-1} r_{i-1}) - P_{i+1} r_{i+1}.  Grid (ceil(n/2), ncols).
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_cr_fwd(const VkCol* __restrict__ cols, int ncols,
                                                   const double* __restrict__ rhs,
                                                   double* __restrict__ cr, int level,
                                                   const int* __restrict__ active) {
  constexpr int BLK = vk_blk<SL>(), N = VK_G * SL;
  __shared__ double rv[3][N];
  const int c = blockIdx.y, w = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const CrLvl v = cr_level(col.nz, BLK, N, 7 * col.off_b, level);
  if (v.n <= 1 || 2 * w >= v.n) return;
  const CrLvl vn = cr_level(col.nz, BLK, N, 7 * col.off_b, level + 1);
  cr_fwd_row<SL>(col, v, vn, level, w, rhs, cr, rv);
}

// x of the top row; for nz == 1 the whole solution.  Grid (ncols).
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_cr_top_solve(const VkCol* __restrict__ cols, int ncols,
                                                         const double* __restrict__ rhs,
                                                         double* __restrict__ cr,
                                                         double* __restrict__ x, SolveEpi epi,
                                                         const int* __restrict__ active) {
  constexpr int BLK = vk_blk<SL>(), N = VK_G * SL;
  __shared__ double rv[N];
  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const int top = cr_top_level(col.nz);
  const CrLvl v = cr_level(col.nz, BLK, N, 7 * col.off_b, top);
  const int ni = col.ni, tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  for (int q = tid; q < N; q += VK_NT)
    rv[q] = q >= ni ? 0.0 : (top == 0 ? rhs[col.off_s + q] : cr[v.r + q]);
  __syncthreads();
  double a[SL][SL], y[SL];
  load_tile<SL>(a, cr + v.Dinv);
  tile_matvec<SL>(a, rv, y);
  if (tc == 0) {
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int q = s * VK_G + tr;
      if (top == 0) {
        if (q < ni) cr_epilogue(epi, col.off_s + q, y[s], x);
      } else {
        cr[v.r + q] = q < ni ? y[s] : 0.0;
      }
    }
  }
}

// Substitute back at one level: even rows copy x from the level above, odd
// rows solve x_j = inv(D_j) ((r_j - L_j x_{j-1}) - U_j x_{j+1}).  Level 0
// writes x and the epilogue.  Grid (n, ncols).
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_cr_bwd(const VkCol* __restrict__ cols, int ncols,
                                                   const double* __restrict__ rhs,
                                                   const double* __restrict__ sup_d,
                                                   const double* __restrict__ sub_d,
                                                   double* __restrict__ cr, int level,
                                                   double* __restrict__ x, SolveEpi epi,
                                                   const int* __restrict__ active) {
  constexpr int BLK = vk_blk<SL>(), N = VK_G * SL;
  __shared__ double xv[3][N];
  const int c = blockIdx.y, j = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const CrLvl v = cr_level(col.nz, BLK, N, 7 * col.off_b, level);
  if (v.n <= 1 || j >= v.n) return;
  const CrLvl vn = cr_level(col.nz, BLK, N, 7 * col.off_b, level + 1);
  const int ni = col.ni, tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  double res[SL];
  if ((j & 1) == 0) {
#pragma unroll
    for (int s = 0; s < SL; ++s) res[s] = cr[vn.r + (int64_t)(j / 2) * N + s * VK_G + tr];
  } else {
    const bool has_u = j + 1 < v.n;
    // xv[0] = x_{j-1}, xv[1] = r_j, xv[2] = x_{j+1} (level-l indices)
    for (int q = tid; q < N; q += VK_NT) {
      const bool in = q < ni;
      xv[0][q] = in ? cr[vn.r + (int64_t)((j - 1) / 2) * N + q] : 0.0;
      xv[2][q] = (in && has_u) ? cr[vn.r + (int64_t)((j + 1) / 2) * N + q] : 0.0;
      xv[1][q] = !in ? 0.0
                     : (level == 0 ? rhs[col.off_s + (int64_t)j * ni + q]
                                   : cr[v.r + (int64_t)j * N + q]);
    }
    __syncthreads();
    double t[SL], y[SL], a[SL][SL];
#pragma unroll
    for (int s = 0; s < SL; ++s) t[s] = 0.0;
    if (level == 0) {
      // L_j = diag(sub_d[j-1]), U_j = diag(sup_d[j])
      if (tc == 0) {
#pragma unroll
        for (int s = 0; s < SL; ++s) {
          const int q = s * VK_G + tr;
          double r = 0.0;
          if (q < ni) {
            r = xv[1][q] - sub_d[col.off_s + (int64_t)(j - 1) * ni + q] * xv[0][q];
            if (has_u) r = r - sup_d[col.off_s + (int64_t)j * ni + q] * xv[2][q];
          }
          t[s] = r;
        }
      }
    } else {
#pragma unroll
      for (int s = 0; s < SL; ++s) t[s] = xv[1][s * VK_G + tr];
      load_tile<SL>(a, cr + v.L + (int64_t)j * BLK);
      tile_matvec<SL>(a, xv[0], y);
#pragma unroll
      for (int s = 0; s < SL; ++s) t[s] = t[s] - y[s];
      if (has_u) {
        load_tile<SL>(a, cr + v.U + (int64_t)j * BLK);
        tile_matvec<SL>(a, xv[2], y);
#pragma unroll
        for (int s = 0; s < SL; ++s) t[s] = t[s] - y[s];
      }
    }
    __syncthreads();
    if (tc == 0) {
#pragma unroll
      for (int s = 0; s < SL; ++s) xv[1][s * VK_G + tr] = t[s];
    }
    __syncthreads();
    load_tile<SL>(a, cr + v.Dinv + (int64_t)((j - 1) / 2) * BLK);
    tile_matvec<SL>(a, xv[1], res);
  }
  if (tc == 0) {
#pragma unroll
    for (int s = 0; s < SL; ++s) {
      const int q = s * VK_G + tr;
      if (level == 0) {
        if (q < ni) cr_epilogue(epi, col.off_s + (int64_t)j * ni + q, res[s], x);
      } else {
        cr[v.r + (int64_t)j * N + q] = q < ni ? res[s] : 0.0;
      }
    }
  }
}

// Natural (nz, ni, ni) blocks, or their transposes, into the tile layout the
// cyclic-reduction factor reads; zero in the padding.  Grid (nz, ncols).
template <int SL>
__global__ __launch_bounds__(VK_NT) void k_nat_to_tile(const VkCol* __restrict__ cols, int ncols,
                                                       const double* __restrict__ nat,
                                                       double* __restrict__ tile, int transpose) {
  const int c = blockIdx.y;
  const int j = blockIdx.x;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL || j >= col.nz) return;
  const int ni = col.ni, tid = threadIdx.x, tr = tid >> 4, tc = tid & 15;
  const double* src = nat + col.off_s * ni + (int64_t)j * ni * ni;
  double* dst = tile + col.off_b + (int64_t)j * vk_blk<SL>();
  for (int s = 0; s < SL; ++s)
    for (int t = 0; t < SL; ++t)
      dst[(s * SL + t) * VK_NT + tid] = nat_elem(src, ni, s * VK_G + tr, t * VK_G + tc, transpose);
}

// ---------------------------------------------------------------------------
// host launchers (dispatch on the slot count of the batch bucket)
// ---------------------------------------------------------------------------

#define VK_DISPATCH_SL(sl, STMT)                 \
  switch (sl) {                                  \
    case 1: { constexpr int SL_ = 1; STMT; } break; \
    case 2: { constexpr int SL_ = 2; STMT; } break; \
    case 3: { constexpr int SL_ = 3; STMT; } break; \
    case 4: { constexpr int SL_ = 4; STMT; } break; \
    case 5: { constexpr int SL_ = 5; STMT; } break; \
    case 6: { constexpr int SL_ = 6; STMT; } break; \
    case 7: { constexpr int SL_ = 7; STMT; } break; \
    default: return VK_ERR_UNSUPPORTED;          \
  }

int vk_launch_bt_factor_rows(int sl, const VkCol* cols, int ncols, const double* diag,
                             int diag_tile, int transpose, const double* sup_d,
                             const double* sub_d, double* inv, const double* rhs, double* rmod,
                             int* status, const int* active, hipStream_t stream);

int vk_launch_bt_factor2(int sl, int nt, const VkCol* cols, int ncols, const double* diag,
                         int diag_tile, int transpose, const double* sup_d, const double* sub_d,
                         double* inv, const double* rhs, double* rmod, int* status,
                         const int* active, hipStream_t stream);
int vk_launch_bt_factor_blk(int sl, const VkCol* cols, int ncols, const double* diag,
                            int diag_tile, int transpose, const double* sup_d, const double* sub_d,
                            double* inv, int* status, const int* active, hipStream_t stream);

// Factor variant per SL bucket: VK_FACTOR_TILE (256 threads, 16x16 register
// tiles), VK_FACTOR_ROWS (half a row per lane), VK_FACTOR_BLOCKED (16-column
// panels, MFMA trailing updates; no fused forward sweep), VK_FACTOR_TILE2,
// VK_FACTOR_WAVE or VK_FACTOR_TILE8 (the tile kernel on 128, 64 or 512
// threads, vk_bt_tile2.hip).  Set by vk_tune_set.
int g_factor_variant[8] = {VK_FACTOR_TILE, VK_FACTOR_TILE, VK_FACTOR_TILE, VK_FACTOR_TILE,
                           VK_FACTOR_TILE, VK_FACTOR_TILE, VK_FACTOR_TILE, VK_FACTOR_TILE};

// Refined sweeps on stored Schur complements: 1 (default) keeps X_j in
// registers for both of its products (k_bt_solve<SL, 3>), 0 reads it twice
// (k_bt_solve<SL, 2>).  Set by vk_tune_set (VK_TUNE_HOLDX).
int g_solve_hold = 1;

bool vk_factor_stores_schur(int sl) {
  return sl >= 1 && sl <= 7 && g_factor_variant[sl] == VK_FACTOR_TILE;
}

bool vk_factor_fuses_sweep(int sl) {
  return !(sl >= 1 && sl <= 7 && (g_factor_variant[sl] == VK_FACTOR_BLOCKED ||
                                  g_factor_variant[sl] == VK_FACTOR_CR));
}

int vk_launch_bt_factor(int sl, const VkCol* cols, int ncols, const double* diag, int diag_tile,
                        int transpose, const double* sup_d, const double* sub_d, double* inv,
                        const double* rhs, double* rmod, int* status, const int* active,
                        hipStream_t stream, double* schur) {
  if (schur && !vk_factor_stores_schur(sl)) {
    vk_set_error("block factor: the selected variant does not store the Schur complements");
    return VK_ERR_UNSUPPORTED;
  }
  if (sl >= 1 && sl <= 7 &&
      (g_factor_variant[sl] == VK_FACTOR_TILE2 || g_factor_variant[sl] == VK_FACTOR_WAVE ||
       g_factor_variant[sl] == VK_FACTOR_TILE8)) {
    const int nt = g_factor_variant[sl] == VK_FACTOR_TILE2 ? 128
                   : g_factor_variant[sl] == VK_FACTOR_WAVE ? 64 : 512;
    return vk_launch_bt_factor2(sl, nt, cols, ncols, diag, diag_tile, transpose, sup_d, sub_d,
                                inv, rhs, rmod, status, active, stream);
  }
  if (sl >= 1 && sl <= 6 && g_factor_variant[sl] == VK_FACTOR_ROWS)
    return vk_launch_bt_factor_rows(sl, cols, ncols, diag, diag_tile, transpose, sup_d, sub_d,
                                    inv, rhs, rmod, status, active, stream);
  if (sl >= 1 && sl <= 7 && g_factor_variant[sl] == VK_FACTOR_BLOCKED) {
    if (rhs) {
      vk_set_error("blocked block factor: the forward sweep is not fused");
      return VK_ERR_UNSUPPORTED;
    }
    return vk_launch_bt_factor_blk(sl, cols, ncols, diag, diag_tile, transpose, sup_d, sub_d, inv,
                                   status, active, stream);
  }
  if (diag_tile) {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_bt_factor<SL_, true>), dim3(ncols), dim3(VK_NT), 0,
                                          stream, cols, ncols, diag, transpose, sup_d, sub_d,
                                          inv, rhs, rmod, status, active, schur));
  } else {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_bt_factor<SL_, false>), dim3(ncols), dim3(VK_NT), 0,
                                          stream, cols, ncols, diag, transpose, sup_d, sub_d,
                                          inv, rhs, rmod, status, active, schur));
  }
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_bt_solve(int sl, const VkCol* cols, int ncols, const double* inv,
                       const double* sup_d, const double* sub_d, const double* rhs,
                       double* rscratch, double* x, int mode, SolveEpi epi, const int* active,
                       hipStream_t stream, const double* refine_diag, int* refl,
                       const double* refine_schur) {
  if (refine_schur && g_solve_hold) {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_bt_solve<SL_, 3>), dim3(ncols), dim3(VK_NT), 0,
                                          stream, cols, ncols, inv, sup_d, sub_d, rhs, rscratch,
                                          x, mode, epi, active, refine_schur, refl));
  } else if (refine_schur) {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_bt_solve<SL_, 2>), dim3(ncols), dim3(VK_NT), 0,
                                          stream, cols, ncols, inv, sup_d, sub_d, rhs, rscratch,
                                          x, mode, epi, active, refine_schur, refl));
  } else if (refine_diag) {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_bt_solve<SL_, 1>), dim3(ncols), dim3(VK_NT), 0,
                                          stream, cols, ncols, inv, sup_d, sub_d, rhs, rscratch,
                                          x, mode, epi, active, refine_diag, refl));
  } else {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_bt_solve<SL_, 0>), dim3(ncols), dim3(VK_NT), 0,
                                          stream, cols, ncols, inv, sup_d, sub_d, rhs, rscratch,
                                          x, mode, epi, active, nullptr, nullptr));
  }
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_bt_residual(int sl, const VkCol* cols, int ncols, int max_nz, const double* blk,
                          const double* sup_d, const double* sub_d, const double* rhs,
                          const double* x, double* res, const int* active, hipStream_t stream) {
  VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_bt_residual<SL_>, dim3(max_nz, ncols), dim3(VK_NT), 0,
                                        stream, cols, ncols, blk, sup_d, sub_d, rhs, x, res,
                                        active));
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_bt_dense(int sl, const VkCol* cols, int ncols, const double* diag,
                       const double* sup, const double* sub, const double* rhs, double* x,
                       double* inv, double* rmod, int* status, hipStream_t stream) {
  VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_bt_dense<SL_>, dim3(ncols), dim3(VK_NT), 0, stream,
                                        cols, ncols, diag, sup, sub, rhs, x, inv, rmod,
                                        status));
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_tile_to_nat(int sl, const VkCol* cols, int ncols, int max_nz, const double* tile,
                          double* nat, hipStream_t stream) {
  VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_tile_to_nat<SL_>, dim3(max_nz, ncols), dim3(VK_NT), 0,
                                        stream, cols, ncols, tile, nat));
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_nat_to_tile(int sl, const VkCol* cols, int ncols, int max_nz, const double* nat,
                          double* tile, int transpose, hipStream_t stream) {
  VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_nat_to_tile<SL_>, dim3(max_nz, ncols), dim3(VK_NT), 0,
                                        stream, cols, ncols, nat, tile, transpose));
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

bool vk_factor_is_cr(int sl) {
  return sl >= 1 && sl <= 7 && g_factor_variant[sl] == VK_FACTOR_CR;
}

// Cyclic-reduction factor of the tile-layout diagonal stack blk (level 0)
// into the region cr (7 * block_doubles); one launch pair per level.  With
// rhs_fused the update launches also reduce that right-hand side, and the
// solve for it is called with fwd_done.
// Levels whose launch has at least this many workgroups (one per block to
// invert, idle ones of other columns included) take the four-per-CU
// elimination and update kernels: three or more per CU of the 256 CUs.
constexpr int64_t kCrWideBlocks = 768;

int vk_launch_cr_factor(int sl, const VkCol* cols, int ncols, int max_nz, const double* blk,
                        const double* sup_d, const double* sub_d, double* cr, int* status,
                        const int* active, hipStream_t stream, const double* rhs_fused) {
  int level = 0;
  for (int n = max_nz; n > 1; n = (n + 1) / 2, ++level) {
    if ((int64_t)(n / 2) * ncols >= kCrWideBlocks) {
      VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_cr_elim<SL_, true>), dim3(n / 2, ncols),
                                            dim3(VK_NT), 0, stream, cols, ncols, blk, sup_d,
                                            sub_d, cr, level, status, active));
    } else {
      VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_cr_elim<SL_, false>), dim3(n / 2, ncols),
                                            dim3(VK_NT), 0, stream, cols, ncols, blk, sup_d,
                                            sub_d, cr, level, status, active));
    }
    VK_CHECK_HIP(hipGetLastError());
    if ((int64_t)((n + 1) / 2) * ncols >= kCrWideBlocks) {
      VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_cr_update<SL_, true>), dim3((n + 1) / 2, ncols),
                                            dim3(VK_NT), 0, stream, cols, ncols, blk, sup_d,
                                            sub_d, cr, level, rhs_fused, active));
    } else {
      VK_DISPATCH_SL(sl, hipLaunchKernelGGL((k_cr_update<SL_, false>), dim3((n + 1) / 2, ncols),
                                            dim3(VK_NT), 0, stream, cols, ncols, blk, sup_d,
                                            sub_d, cr, level, rhs_fused, active));
    }
    VK_CHECK_HIP(hipGetLastError());
  }
  VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_cr_top<SL_>, dim3(ncols), dim3(VK_NT), 0, stream, cols,
                                        ncols, blk, cr, status, active));
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

// Solve with the cyclic-reduction factors: x and the epilogue from rhs.
int vk_launch_cr_solve(int sl, const VkCol* cols, int ncols, int max_nz, const double* rhs,
                       const double* sup_d, const double* sub_d, double* cr, double* x,
                       SolveEpi epi, const int* active, hipStream_t stream, bool fwd_done) {
  int ns[40];
  int levels = 0;
  for (int n = max_nz; n > 1 && levels < 40; n = (n + 1) / 2) ns[levels++] = n;
  for (int l = 0; l < levels && !fwd_done; ++l) {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_cr_fwd<SL_>, dim3((ns[l] + 1) / 2, ncols),
                                          dim3(VK_NT), 0, stream, cols, ncols, rhs, cr, l,
                                          active));
    VK_CHECK_HIP(hipGetLastError());
  }
  VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_cr_top_solve<SL_>, dim3(ncols), dim3(VK_NT), 0, stream,
                                        cols, ncols, rhs, cr, x, epi, active));
  VK_CHECK_HIP(hipGetLastError());
  for (int l = levels - 1; l >= 0; --l) {
    VK_DISPATCH_SL(sl, hipLaunchKernelGGL(k_cr_bwd<SL_>, dim3(ns[l], ncols), dim3(VK_NT), 0,
                                          stream, cols, ncols, rhs, sup_d, sub_d, cr, l, x, epi,
                                          active));
    VK_CHECK_HIP(hipGetLastError());
  }
  return VK_OK;
}
