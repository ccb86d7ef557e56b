// Block-Thomas factor, half-row-per-lane variant.
//
// Same algebra and outputs as k_bt_factor (vk_bt.hip): per layer the Schur
// update A'_j = D_j - (c_j b_{j-1}^T) o inv(A'_{j-1}), the fused forward
// sweep r'_j = r_j - c_j o (inv(A'_{j-1}) r'_{j-1}), and the in-place
// swap-free Gauss-Jordan inversion with first-max partial pivoting, whose
// result R gives inv(A'_j)[a][b] = R[pi(a)][pi^-1(b)].  Per element the
// arithmetic is the tile kernel's, so the inverse stacks are bitwise equal;
// the inverse is written in the same tile layout and the solves are shared.
//
// Mapping: lanes l and l+32 of a wave hold the two halves (HB column blocks
// of 16 each) of row 32*wave + (l & 31), in registers.  Per pivot step:
//   * the pivot-column value is read from a 16-column register window (the
//     column block being eliminated), so every register index is static;
//     the half that does not own the block gets it with one lane swap;
//   * the first-max row is a wave max plus a ballot (smallest lane of the
//     maximum) and, across waves, one LDS exchange before the one barrier;
//   * each wave's candidate row is published to LDS before that barrier, so
//     the winner is read by every lane as a broadcast;
//   * every lane updates its half row: 16*HB + 16 FMAs.
// The permutation is undone once per layer through a column-major LDS
// buffer.  About N*N*8 bytes of LDS per column: two columns of ni <= 96 per CU.
#include <utility>

#include "vk_common.h"

namespace {

template <int SL>
struct RowsCfg {
  static constexpr int N = VK_G * SL;          // padded rows / columns of the block
  static constexpr int HB = (SL + 1) / 2;      // column blocks per half row
  static constexpr int NL = VK_G * HB;         // columns per lane
  static constexpr int NG = 2 * NL;            // padded columns of the two halves (>= N)
  static constexpr int NWAV = (N + 31) / 32;   // 32 rows per wave
  static constexpr int NT = 64 * NWAV;
  // LDS carve (doubles): [buf | cand] [cval] [vec x2] [bvec], then ints
  static constexpr int CAND = 2 * NWAV * NG;
  static constexpr int BUF = N * N > CAND ? N * N : CAND;
  static constexpr int OFF_CVAL = BUF;
  static constexpr int OFF_VEC = OFF_CVAL + 2 * NWAV;
  static constexpr int OFF_BVEC = OFF_VEC + 2 * NG;
  static constexpr int NDBL = OFF_BVEC + NG;
  static constexpr size_t BYTES = sizeof(double) * NDBL + sizeof(int) * (2 * NWAV + 2 * N);
};

// Lane's half row from tile layout (row r, global blocks h*HB .. h*HB+HB-1).
template <int SL>
__device__ __forceinline__ void load_half_tile(double (&row)[RowsCfg<SL>::NL], const double* blk,
                                               int r, int h) {
  constexpr int HB = RowsCfg<SL>::HB;
  const int sa = r >> 4, ra = r & 15;
#pragma unroll
  for (int lb = 0; lb < HB; ++lb) {
    const int b = h * HB + lb;
    if (b < SL) {
      const double2* src = (const double2*)(blk + (sa * SL + b) * VK_NT + ra * VK_G);
#pragma unroll
      for (int c = 0; c < VK_G / 2; ++c) {
        const double2 v = src[c];
        row[lb * VK_G + 2 * c] = v.x;
        row[lb * VK_G + 2 * c + 1] = v.y;
      }
    } else {
#pragma unroll
      for (int c = 0; c < VK_G; ++c) row[lb * VK_G + c] = 0.0;
    }
  }
}

template <int SL>
__device__ __forceinline__ void store_half_tile(const double (&row)[RowsCfg<SL>::NL], double* blk,
                                                int r, int h) {
  constexpr int HB = RowsCfg<SL>::HB;
  const int sa = r >> 4, ra = r & 15;
#pragma unroll
  for (int lb = 0; lb < HB; ++lb) {
    const int b = h * HB + lb;
    if (b < SL) {
      double2* dst = (double2*)(blk + (sa * SL + b) * VK_NT + ra * VK_G);
#pragma unroll
      for (int c = 0; c < VK_G / 2; ++c)
        dst[c] = make_double2(row[lb * VK_G + 2 * c], row[lb * VK_G + 2 * c + 1]);
    }
  }
}

// Stage a natural (ni, ni) block (or its transpose) into the column-major
// N x N LDS buffer, zero in the padding; barriers on both sides.
template <int SL>
__device__ __forceinline__ void stage_natural(double* buf, const double* blk, int ni, int transpose) {
  constexpr int N = RowsCfg<SL>::N;
  __syncthreads();
  for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
    const int r = e / N, cc = e - r * N;
    double v = 0.0;
    if (r < ni && cc < ni) v = transpose ? blk[(int64_t)cc * ni + r] : blk[(int64_t)r * ni + cc];
    buf[cc * N + r] = v;
  }
  __syncthreads();
}

// The 16-column window of local block lb, blended from the static blocks
// with bit masks (a select or a switch would be folded into one dynamically
// indexed access, which sends the row to scratch).  lb outside [0, HB):
// the window reads as zero and is not stored back.
template <int SL>
__device__ __forceinline__ void window_load(const double (&row)[RowsCfg<SL>::NL], double (&w)[VK_G],
                                            int lb) {
  constexpr int HB = RowsCfg<SL>::HB;
  long long m[HB];
#pragma unroll
  for (int B = 0; B < HB; ++B) m[B] = -(long long)(lb == B);
#pragma unroll
  for (int t = 0; t < VK_G; ++t) {
    long long v = 0;
#pragma unroll
    for (int B = 0; B < HB; ++B) v |= __double_as_longlong(row[B * VK_G + t]) & m[B];
    w[t] = __longlong_as_double(v);
  }
}

template <int SL>
__device__ __forceinline__ void window_store(double (&row)[RowsCfg<SL>::NL], const double (&w)[VK_G],
                                             int lb) {
  constexpr int HB = RowsCfg<SL>::HB;
#pragma unroll
  for (int B = 0; B < HB; ++B) {
    const long long mb = -(long long)(lb == B);
#pragma unroll
    for (int t = 0; t < VK_G; ++t)
      row[B * VK_G + t] = __longlong_as_double((__double_as_longlong(w[t]) & mb) |
                                               (__double_as_longlong(row[B * VK_G + t]) & ~mb));
  }
}

// One step of a wave max: v = max(v, v moved by a DPP pattern); lanes with
// no source (or outside ROW_MASK) keep their own value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_max_step(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
  return fmax(v, __hiloint2double(hi2, lo2));
}

// Max over the 64 lanes of a wave (no NaN inputs), returned in every lane:
// row_shr 1/2/4/8 leaves each 16-lane row's max in its lane 15, row_bcast 15
// and 31 carry them to lane 63, which is read back as a scalar.
__device__ __forceinline__ double wave_max(double v) {
  v = dpp_max_step<0x111, 0xF>(v);
  v = dpp_max_step<0x112, 0xF>(v);
  v = dpp_max_step<0x114, 0xF>(v);
  v = dpp_max_step<0x118, 0xF>(v);
  v = dpp_max_step<0x142, 0xA>(v);
  v = dpp_max_step<0x143, 0xC>(v);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

struct StepCtx {
  double* smem;
  int n, r, h, kb, hk;  // hk: the half that owns block kb
  bool used;
  int bad;
};

// One pivot step, pivot column k = 16*kb + KC (KC static: register indices
// stay static).
template <int SL, int KC>
__device__ __forceinline__ void gj_step(double (&row)[RowsCfg<SL>::NL], double (&w)[VK_G],
                                        StepCtx& s) {
  using Cf = RowsCfg<SL>;
  constexpr int NL = Cf::NL, NG = Cf::NG, NWAV = Cf::NWAV;
  const int k = s.kb * VK_G + KC;
  if (k >= s.n) return;  // uniform: the last block may be partial
  double* cand = s.smem;
  double* cval = s.smem + Cf::OFF_CVAL;
  int* cidx = (int*)(s.smem + Cf::NDBL);
  int* perm = cidx + 2 * NWAV;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int par = k & 1;
  // The half that owns block kb holds column k in its window and runs the
  // search; the other half gets the value from its partner lane.
  const bool owner = (s.h == s.hk);
  const double vo = w[KC];
  const double vx = __shfl_xor(vo, 32, 64);
  const bool live = owner && (s.r < s.n) && !s.used;
  double av = fabs(vo);
  if (!(av >= 0.0)) av = -0.5;  // NaN: below every number, above "none"
  const double cand_v = live ? av : -1.0;
  const double m = wave_max(cand_v);
  const unsigned long long bal = __ballot(live && cand_v == m);
  const int fl = bal ? (__builtin_ctzll(bal) & 31) : 64;  // row within the wave
  double* myc = cand + (par * NWAV + wave) * NG;
  if (bal && (lane & 31) == fl) {
#pragma unroll
    for (int q = 0; q < NL; ++q) myc[s.h * NL + q] = row[q];
    if (owner) {
#pragma unroll
      for (int t = 0; t < VK_G; ++t) myc[s.kb * VK_G + t] = w[t];
    }
  }
  if (lane == 0) {
    cval[par * NWAV + wave] = bal ? m : -1.0;
    cidx[par * NWAV + wave] = bal ? wave * 32 + fl : 0x7fffffff;
  }
  __syncthreads();
  double pv = cval[par * NWAV];
  int p = cidx[par * NWAV], pw = 0;
#pragma unroll
  for (int ww = 1; ww < NWAV; ++ww) {
    const double ov = cval[par * NWAV + ww];
    const int oi = cidx[par * NWAV + ww];
    if (ov > pv || (ov == pv && oi < p)) { pv = ov; p = oi; pw = ww; }
  }
  p = __builtin_amdgcn_readfirstlane(p);
  pw = __builtin_amdgcn_readfirstlane(pw);
  if ((unsigned)p >= (unsigned)s.n) {  // cannot happen while k < n; keeps indices in range
    p = k;
    s.bad = 1;
  }
  const double* u = cand + (par * NWAV + pw) * NG;
  const double piv = u[k];
  if (!(piv != 0.0) || !isfinite(piv)) s.bad = 1;
  const double rpiv = 1.0 / piv;
  if (threadIdx.x == 0) perm[k] = p;
  const double v = owner ? vo : vx;  // this row's entry in column k
  const double l = v * rpiv;
  // Every row but the pivot row: a[i][.] -= l_i u, column k := -l_i.  The
  // pivot row equals u (it published it), so it skips the update (sc = 0)
  // and is scaled in place: a[p][.] = u * rpiv, a[p][k] = rpiv.
  const bool is_p = (s.r == p);
  const double sc = is_p ? 0.0 : -l;
  const double* uh = u + s.h * NL;
#pragma unroll
  for (int q = 0; q < NL; ++q) row[q] = fma(sc, uh[q], row[q]);
  const double* uw = u + s.kb * VK_G;
#pragma unroll
  for (int t = 0; t < VK_G; ++t) w[t] = fma(sc, uw[t], w[t]);
  w[KC] = -l;
  if (is_p) {
#pragma unroll
    for (int q = 0; q < NL; ++q) row[q] = row[q] * rpiv;
#pragma unroll
    for (int t = 0; t < VK_G; ++t) w[t] = w[t] * rpiv;
    w[KC] = rpiv;
    s.used = true;
  }
}

template <int SL, int... KC>
__device__ __forceinline__ void gj_block(double (&row)[RowsCfg<SL>::NL], double (&w)[VK_G],
                                         StepCtx& s, std::integer_sequence<int, KC...>) {
  (gj_step<SL, KC>(row, w, s), ...);
}

}  // namespace

template <int SL>
__global__ __launch_bounds__(RowsCfg<SL>::NT) void k_bt_factor_rows(
    const VkCol* __restrict__ cols, int ncols, const double* diag, int diag_tile, int transpose,
    const double* __restrict__ sup_d, const double* __restrict__ sub_d, double* inv,
    const double* __restrict__ rhs, double* __restrict__ rmod, int* __restrict__ status,
    const int* __restrict__ active) {
  using Cf = RowsCfg<SL>;
  constexpr int N = Cf::N, NL = Cf::NL, NG = Cf::NG, NWAV = Cf::NWAV, HB = Cf::HB;
  extern __shared__ double smem[];
  double* buf = smem;                       // N x N column-major staging (aliases cand)
  double* vec = smem + Cf::OFF_VEC;         // [2][NG]
  double* bvec = smem + Cf::OFF_BVEC;       // [NG]
  int* cidx = (int*)(smem + Cf::NDBL);
  int* perm = cidx + 2 * NWAV;              // [N]
  int* iperm = perm + N;                    // [N]

  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const int nz = col.nz, n = col.ni;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r = wave * 32 + (lane & 31), h = lane >> 5;
  const int q0 = h * NL;  // first global column of this lane's half
  constexpr int BLK = vk_blk<SL>();
  const double* bsup = sup_d + col.off_s;
  const double* bsub = sub_d + col.off_s;
  const bool sweep = (rhs != nullptr);
  const bool own = r < N;

  double row[NL];
#pragma unroll
  for (int q = 0; q < NL; ++q) row[q] = 0.0;
  StepCtx s{smem, n, r, h, 0, 0, false, 0};

  for (int j = 0; j < nz; ++j) {
    const double* Dj_nat = diag + col.off_s * n + (int64_t)j * n * n;
    const double* Dj_tile = diag + col.off_b + (int64_t)j * BLK;
    if (!diag_tile) stage_natural<SL>(buf, Dj_nat, n, transpose);
    if (j == 0) {
      if (own) {
        if (diag_tile) {
          load_half_tile<SL>(row, Dj_tile, r, h);
        } else {
#pragma unroll
          for (int q = 0; q < NL; ++q) row[q] = (q0 + q < N) ? buf[(q0 + q) * N + r] : 0.0;
        }
      }
      if (sweep && own && h == 0) {
        const double v = (r < n) ? rhs[col.off_s + r] : 0.0;
        vec[r] = v;
        if (r < n) rmod[col.off_s + r] = v;
      }
      for (int i = N + tid; i < NG; i += Cf::NT) vec[i] = 0.0;  // padding of both buffers
      for (int i = N + tid; i < NG; i += Cf::NT) vec[NG + i] = 0.0;
    } else {
      for (int i = tid; i < NG; i += Cf::NT) bvec[i] = (i < n) ? bsup[(int64_t)(j - 1) * n + i] : 0.0;
      __syncthreads();
      const double cv = (r < n) ? bsub[(int64_t)(j - 1) * n + r] : 0.0;
      if (sweep) {
        // r'_j = r_j - c_j o (X_{j-1} r'_{j-1}) while row still holds X_{j-1}
        const double* vp = vec + ((j - 1) & 1) * NG + q0;
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < NL; ++q) acc = fma(row[q], vp[q], acc);
        const double other = __shfl_xor(acc, 32, 64);
        acc = h == 0 ? acc + other : other + acc;
        if (own && h == 0) {
          double rr = 0.0;
          if (r < n) {
            rr = rhs[col.off_s + (int64_t)j * n + r] - cv * acc;
            rmod[col.off_s + (int64_t)j * n + r] = rr;
          }
          vec[(j & 1) * NG + r] = rr;
        }
      }
      if (own) {
        const double* bv = bvec + q0;
        if (diag_tile) {
          const int sa = r >> 4, ra = r & 15;
#pragma unroll
          for (int lb = 0; lb < HB; ++lb) {
            const int b = h * HB + lb;
            if (b < SL) {
              const double2* src = (const double2*)(Dj_tile + (sa * SL + b) * VK_NT + ra * VK_G);
#pragma unroll
              for (int cc = 0; cc < VK_G / 2; ++cc) {
                const double2 d = src[cc];
                const int q = lb * VK_G + 2 * cc;
                row[q] = d.x - (cv * bv[q]) * row[q];
                row[q + 1] = d.y - (cv * bv[q + 1]) * row[q + 1];
              }
            }
          }
        } else {
#pragma unroll
          for (int q = 0; q < NL; ++q) {
            const double d = (q0 + q < N) ? buf[(q0 + q) * N + r] : 0.0;
            row[q] = d - (cv * bv[q]) * row[q];
          }
        }
      }
    }

    // ---- in-place Gauss-Jordan inversion with first-max pivoting ----------
    if (!diag_tile) __syncthreads();  // staged block read; cand reuses buf
    s.used = false;
    for (int kb = 0; kb * VK_G < n; ++kb) {
      s.kb = kb;
      s.hk = kb / HB;
      const int lb = kb - h * HB;
      double w[VK_G];
      window_load<SL>(row, w, lb);
      gj_block<SL>(row, w, s, std::make_integer_sequence<int, VK_G>{});
      window_store<SL>(row, w, lb);
    }

    // ---- undo the permutation: X[iperm[r]][perm[q]] = R[r][q] --------------
    __syncthreads();  // candidate rows are no longer read; perm is complete
    for (int i = n + tid; i < N; i += Cf::NT) perm[i] = i;
    __syncthreads();
    if (own && h == 0) iperm[perm[r]] = r;
    __syncthreads();
    if (own) {
      const int dst = iperm[r];
#pragma unroll
      for (int q = 0; q < NL; ++q)
        if (q0 + q < N) buf[perm[q0 + q] * N + dst] = row[q];
    }
    __syncthreads();
    if (own) {
#pragma unroll
      for (int q = 0; q < NL; ++q) row[q] = (q0 + q < N) ? buf[(q0 + q) * N + r] : 0.0;
      store_half_tile<SL>(row, inv + col.off_b + (int64_t)j * BLK, r, h);
    }
    __syncthreads();  // buf is reused as candidate storage by the next layer
  }
  if (s.bad && tid == 0 && status) atomicOr(&status[c], VK_ST_SINGULAR);
}

size_t vk_bt_rows_lds_bytes(int sl) {
  switch (sl) {
    case 1: return RowsCfg<1>::BYTES;
    case 2: return RowsCfg<2>::BYTES;
    case 3: return RowsCfg<3>::BYTES;
    case 4: return RowsCfg<4>::BYTES;
    case 5: return RowsCfg<5>::BYTES;
    case 6: return RowsCfg<6>::BYTES;
    default: return 0;
  }
}

int vk_launch_bt_factor_rows(int sl, const VkCol* cols, int ncols, const double* diag,
                             int diag_tile, int transpose, const double* sup_d,
                             const double* sub_d, double* inv, const double* rhs, double* rmod,
                             int* status, const int* active, hipStream_t stream) {
  static bool raised[8] = {false};
  switch (sl) {
#define VK_ROWS(S)                                                                           \
  case S: {                                                                                  \
    const size_t sh = RowsCfg<S>::BYTES;                                                     \
    if (!raised[S]) {                                                                        \
      if (hipFuncSetAttribute((const void*)k_bt_factor_rows<S>,                             \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh) !=        \
          hipSuccess) {                                                                      \
        (void)hipGetLastError();                                                             \
        vk_set_error("block factor: %zu bytes of LDS not available", sh);                    \
        return VK_ERR_UNSUPPORTED;                                                           \
      }                                                                                      \
      raised[S] = true;                                                                      \
    }                                                                                        \
    hipLaunchKernelGGL(k_bt_factor_rows<S>, dim3(ncols), dim3(RowsCfg<S>::NT), sh, stream,  \
                       cols, ncols, diag, diag_tile, transpose, sup_d, sub_d, inv, rhs, rmod, \
                       status, active);                                                      \
  } break;
    VK_ROWS(1) VK_ROWS(2) VK_ROWS(3) VK_ROWS(4) VK_ROWS(5) VK_ROWS(6)
#undef VK_ROWS
    default:
      vk_set_error("half-row block factor supports ni <= 96");
      return VK_ERR_UNSUPPORTED;
  }
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
