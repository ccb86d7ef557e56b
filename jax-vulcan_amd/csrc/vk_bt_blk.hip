// Block-Thomas factor with a blocked Gauss-Jordan inversion on FP64 MFMA.
//
// Same outputs as k_bt_factor (vk_bt.hip): the stack of inverses inv(A'_j),
// A'_j = D_j - (c_j b_{j-1}^T) o inv(A'_{j-1}), in tile layout, with the
// first-max partial pivoting of the unblocked swap-free Gauss-Jordan
// elimination.  The elimination runs in 16-column panels:
//
//   * the panel (all rows, 16 columns) is eliminated in LDS step by step with
//     exactly the per-element arithmetic of the unblocked kernel;
//   * the other columns receive the panel's 16 row operations at once.  If
//     M = E_15 ... E_0 is the product of the step transforms and P selects
//     the 16 pivot rows, the eliminated panel holds G = M P, and with R = the
//     pivot rows of the other columns before the panel,
//         A[i, rest] += G[i, :] R    for every row i that is not a pivot row,
//         A[i, rest]  = G[i, :] R    for the pivot rows,
//     a rank-16 product on v_mfma_f64_16x16x4f64 over 16x16 register tiles
//     (pivot-row accumulators are zeroed first).  Writing M = I + (G - P) P^T
//     instead would subtract 1 from the pivot scalings 1/piv and lose them
//     whenever |piv| >> 1.
//
// Rounding differs from the unblocked kernel only in that rank-16 sum.  The
// forward sweep is not fused here: the caller runs the full solve.
//
// Mapping: one 256-thread workgroup (4 waves) per column.  Tile (a, b) of the
// padded N x N block (N = 16 SL) is held by wave (a SL + b) mod 4 in MFMA
// result layout: lane l, item v holds element (16a + l/16 + 4v, 16b + l%16).
#include "vk_common.h"

namespace {

typedef double vk_d4 __attribute__((ext_vector_type(4)));

template <int SL>
struct BlkCfg {
  static constexpr int N = VK_G * SL;
  static constexpr int NTILE = SL * SL;
  static constexpr int TPW = (NTILE + 3) / 4;  // tiles per wave
  // LDS (doubles): panel ping-pong P[2][N][16], C[N][16], R[16][N]; the
  // unpermute buffer [N][N] (column-major) aliases them.
  static constexpr int OFF_P = 0;
  static constexpr int OFF_C = 2 * N * VK_G;
  static constexpr int OFF_R = OFF_C + N * VK_G;
  static constexpr int OFF_B = OFF_R + VK_G * N;  // b (sup) and c (sub) vectors
  static constexpr int PANEL_END = OFF_B + 2 * N;
  static constexpr int NDBL = PANEL_END > N * N + 2 * N ? PANEL_END : N * N + 2 * N;
  static constexpr size_t BYTES = sizeof(double) * NDBL + sizeof(int) * (3 * N + VK_G);
};

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_max_step(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
  return fmax(v, __hiloint2double(hi2, lo2));
}

// Max over the 64 lanes of a wave (no NaN inputs), in every lane.
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

__device__ __forceinline__ double pick4(const vk_d4& t, int v) {
  const double a = v == 1 ? t[1] : t[0];
  const double b = v == 3 ? t[3] : t[2];
  return v >= 2 ? b : a;
}

}  // namespace

template <int SL>
__global__ __launch_bounds__(VK_NT) void k_bt_factor_blk(
    const VkCol* __restrict__ cols, int ncols, const double* diag, int diag_tile, int transpose,
    const double* __restrict__ sup_d, const double* __restrict__ sub_d, double* inv,
    int* __restrict__ status, const int* __restrict__ active) {
  using Cf = BlkCfg<SL>;
  constexpr int N = Cf::N, TPW = Cf::TPW, NTILE = Cf::NTILE;
  extern __shared__ double smem[];
  double* P = smem + Cf::OFF_P;
  double* Cb = smem + Cf::OFF_C;
  double* Rb = smem + Cf::OFF_R;
  double* buf = smem;                      // unpermute staging, column-major (aliases P/C/R)
  double* bvec = smem + Cf::NDBL - 2 * N;  // b_{j-1} (columns) and c_j (rows) of the Schur update
  double* cvec = bvec + N;
  int* perm = (int*)(smem + Cf::NDBL);
  int* iperm = perm + N;
  int* pv = iperm + N;                     // pivot rows of the current panel
  int* isp = pv + VK_G;                    // [N] row is a pivot row of the current panel

  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (vk_slots(col.ni) != SL) return;
  const int nz = col.nz, n = col.ni;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr0 = lane >> 4, lc0 = lane & 15;  // row offset (+4v) and column within a tile
  constexpr int BLK = vk_blk<SL>();
  const double* bsup = sup_d + col.off_s;
  const double* bsub = sub_d + col.off_s;

  vk_d4 T[TPW];
#pragma unroll
  for (int s = 0; s < TPW; ++s) T[s] = vk_d4{0.0, 0.0, 0.0, 0.0};
  int bad = 0;
  // rows l and l + 64 of the pivot search (the same in every wave)
  const int i0 = lane, i1 = lane + 64;

  for (int j = 0; j < nz; ++j) {
    // Lane-derived indices are redefined every layer (and panel) so the
    // compiler cannot hoist dozens of per-element addresses out of the loops
    // and run out of registers.
    int lr = lr0, lc = lc0;
    asm volatile("" : "+v"(lr), "+v"(lc));
    const double* Dj_nat = diag + col.off_s * n + (int64_t)j * n * n;
    const double* Dj_tile = diag + col.off_b + (int64_t)j * BLK;
    if (j > 0) {
      for (int i = tid; i < N; i += VK_NT) {
        bvec[i] = i < n ? bsup[(int64_t)(j - 1) * n + i] : 0.0;
        cvec[i] = i < n ? bsub[(int64_t)(j - 1) * n + i] : 0.0;
      }
      __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
      const int t = wave + 4 * s;
      if (t < NTILE) {
        const int a = t / SL, b = t - a * SL;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = lr + 4 * v, ii = 16 * a + r, jj = 16 * b + lc;
          double d;
          if (diag_tile) {
            d = Dj_tile[(a * SL + b) * VK_NT + r * VK_G + lc];
          } else {
            d = (ii < n && jj < n)
                    ? (transpose ? Dj_nat[(int64_t)jj * n + ii] : Dj_nat[(int64_t)ii * n + jj])
                    : 0.0;
          }
          T[s][v] = j == 0 ? d : d - (cvec[ii] * bvec[jj]) * T[s][v];
        }
      }
    }

    // ---- blocked Gauss-Jordan inversion ------------------------------------
    bool used0 = false, used1 = false;
    for (int kb = 0; kb * VK_G < n; ++kb) {
      const int mcount = (n - VK_G * kb) < VK_G ? (n - VK_G * kb) : VK_G;
      asm volatile("" : "+v"(lr), "+v"(lc));
      for (int i = tid; i < N; i += VK_NT) isp[i] = 0;
      // panel tiles -> P[0]
#pragma unroll
      for (int s = 0; s < TPW; ++s) {
        const int t = wave + 4 * s;
        if (t < NTILE && t % SL == kb) {
          const int a = t / SL;
#pragma unroll
          for (int v = 0; v < 4; ++v) P[(16 * a + lr + 4 * v) * VK_G + lc] = T[s][v];
        }
      }
      __syncthreads();
      for (int m = 0; m < mcount; ++m) {
        const int k = VK_G * kb + m;
        const double* cur = P + (m & 1) * N * VK_G;
        double* nxt = P + ((m + 1) & 1) * N * VK_G;
        // first-max pivot over the unused rows of column m (every wave alike)
        double v0 = -1.0, v1 = -1.0;
        if (i0 < n && !used0) {
          v0 = fabs(cur[i0 * VK_G + m]);
          if (!(v0 >= 0.0)) v0 = -0.5;  // NaN: below every number, above "none"
        }
        if (i1 < n && !used1) {
          v1 = fabs(cur[i1 * VK_G + m]);
          if (!(v1 >= 0.0)) v1 = -0.5;
        }
        const bool take1 = v1 > v0;
        const double bv = take1 ? v1 : v0;
        const double mx = wave_max(bv);
        const unsigned long long lo_bal = __ballot(bv == mx && bv > -1.0 && !take1);
        const unsigned long long hi_bal = __ballot(bv == mx && bv > -1.0 && take1);
        int p;
        if (lo_bal) p = __builtin_ctzll(lo_bal);
        else if (hi_bal) p = 64 + __builtin_ctzll(hi_bal);
        else { p = k; bad = 1; }  // cannot happen while k < n; keeps indices in range
        p = __builtin_amdgcn_readfirstlane(p);
        const double piv = cur[p * VK_G + m];
        if (!(piv != 0.0) || !isfinite(piv)) bad = 1;
        const double rpiv = 1.0 / piv;
        for (int e = tid; e < N * VK_G; e += VK_NT) {
          const int i = e >> 4, jj = e & 15;
          const double u = cur[p * VK_G + jj];
          double val;
          if (i == p) {
            val = jj == m ? rpiv : u * rpiv;
          } else {
            const double l = cur[i * VK_G + m] * rpiv;
            val = jj == m ? -l : fma(-l, u, cur[e]);
          }
          nxt[e] = val;
        }
        if (i0 == p) used0 = true;
        if (i1 == p) used1 = true;
        if (tid == 0) { perm[k] = p; pv[m] = p; isp[p] = 1; }
        __syncthreads();
      }
      const double* Pf = P + (mcount & 1) * N * VK_G;
      // G = M P (the eliminated panel), and R = the pivot rows of the other
      // columns before the panel
      for (int e = tid; e < N * VK_G; e += VK_NT) {
        const int m = e & 15;
        Cb[e] = m < mcount ? Pf[e] : 0.0;
      }
      for (int e = tid; e < VK_G * N; e += VK_NT) Rb[e] = 0.0;
      __syncthreads();
#pragma unroll
      for (int s = 0; s < TPW; ++s) {
        const int t = wave + 4 * s;
        if (t < NTILE && t % SL != kb) {
          const int a = t / SL, b = t - a * SL;
          for (int m = 0; m < mcount; ++m) {
            const int pr = pv[m];
            if ((pr >> 4) == a && lr == (pr & 3)) Rb[m * N + 16 * b + lc] = pick4(T[s], (pr & 15) >> 2);
          }
        }
      }
      __syncthreads();
      // panel tiles <- eliminated panel; the others += C R on MFMA
#pragma unroll
      for (int s = 0; s < TPW; ++s) {
        const int t = wave + 4 * s;
        if (t < NTILE) {
          const int a = t / SL, b = t - a * SL;
          if (b == kb) {
#pragma unroll
            for (int v = 0; v < 4; ++v) T[s][v] = Pf[(16 * a + lr + 4 * v) * VK_G + lc];
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v)
              if (isp[16 * a + lr + 4 * v]) T[s][v] = 0.0;  // pivot rows are replaced
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const double av = Cb[(16 * a + lc) * VK_G + 4 * q + lr];
              const double bw = Rb[(4 * q + lr) * N + 16 * b + lc];
              T[s] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bw, T[s], 0, 0, 0);
            }
          }
        }
      }
      __syncthreads();  // P, C, R are rewritten by the next panel
    }

    // ---- undo the permutation: X[iperm[r]][perm[q]] = R[r][q] --------------
    for (int i = n + tid; i < N; i += VK_NT) perm[i] = i;
    __syncthreads();
    for (int i = tid; i < N; i += VK_NT) iperm[perm[i]] = i;
    __syncthreads();
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
      const int t = wave + 4 * s;
      if (t < NTILE) {
        const int a = t / SL, b = t - a * SL;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          buf[perm[16 * b + lc] * N + iperm[16 * a + lr + 4 * v]] = T[s][v];
      }
    }
    __syncthreads();
    double* dst = inv + col.off_b + (int64_t)j * BLK;
#pragma unroll
    for (int s = 0; s < TPW; ++s) {
      const int t = wave + 4 * s;
      if (t < NTILE) {
        const int a = t / SL, b = t - a * SL;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int r = lr + 4 * v;
          T[s][v] = buf[(16 * b + lc) * N + 16 * a + r];
          dst[(a * SL + b) * VK_NT + r * VK_G + lc] = T[s][v];
        }
      }
    }
    __syncthreads();  // buf aliases the panel storage of the next layer
  }
  if (bad && tid == 0 && status) atomicOr(&status[c], VK_ST_SINGULAR);
}

int vk_launch_bt_factor_blk(int sl, const VkCol* cols, int ncols, const double* diag,
                            int diag_tile, int transpose, const double* sup_d, const double* sub_d,
                            double* inv, int* status, const int* active, hipStream_t stream) {
  static bool raised[8] = {false};
  switch (sl) {
#define VK_BLK(S)                                                                            \
  case S: {                                                                                  \
    const size_t sh = BlkCfg<S>::BYTES;                                                      \
    if (!raised[S]) {                                                                        \
      if (hipFuncSetAttribute((const void*)k_bt_factor_blk<S>,                              \
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh) !=        \
          hipSuccess) {                                                                      \
        (void)hipGetLastError();                                                             \
        vk_set_error("blocked block factor: %zu bytes of LDS not available", sh);            \
        return VK_ERR_UNSUPPORTED;                                                           \
      }                                                                                      \
      raised[S] = true;                                                                      \
    }                                                                                        \
    hipLaunchKernelGGL(k_bt_factor_blk<S>, dim3(ncols), dim3(VK_NT), sh, stream, cols, ncols, \
                       diag, diag_tile, transpose, sup_d, sub_d, inv, status, active);       \
  } break;
    VK_BLK(1) VK_BLK(2) VK_BLK(3) VK_BLK(4) VK_BLK(5) VK_BLK(6) VK_BLK(7)
#undef VK_BLK
    default:
      vk_set_error("blocked block factor: unsupported bucket %d", sl);
      return VK_ERR_UNSUPPORTED;
  }
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
