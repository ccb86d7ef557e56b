// Photolysis: optical depth, two-stream Eddington flux, J-rates, k update.
//
// Restates photo.compute_tau_jax (photo.py:84-123), compute_flux_jax
// (photo.py:126-245), _compute_J_inner (photo.py:309-351) and
// update_k_with_J (photo.py:455-478).  All columns of a launch share one
// cross-section / stellar-flux setup; each column has its own state arrays
// of shape (nz+1, nbin) at off_p[c].
//
// k_photo_optics: one thread per wavelength bin, sequential in z.  Per layer
//   the absorber/scatterer densities are staged in LDS and every thread
//   forms its bin's sums; tau is accumulated top-down, w0 kept per layer.
// k_photo_flux: one thread per bin; the down sweep uses the PREVIOUS call's
//   dflux_u (photo.py:141-145, 210-217); the up sweep recomputes the layer
//   coefficients from tau and w0 and forms the actinic flux.
// k_photo_J: the contraction J = cross_J (n_br x nbin) . aflux^T (nbin x nz)
//   on FP64 MFMA, one workgroup per (16-layer tile, column): its waves split
//   the bins and every 16-branch tile of a 64-branch group; the slab sums
//   are added in a fixed order.  The
//   two trapezoid spacings ride in the flux operand (af * dbin), so one
//   accumulator per tile holds both partial sums; the half end-point terms
//   are subtracted after.  J goes to the state and, times f_diurnal, into
//   the k rows.
// k_photo_JT: the temperature-dependent branches, whose cross section
//   differs per layer (a dot product per layer, not a contraction): one
//   workgroup per (layer, column) as the split trapezoid.
#include "vk_common.h"

namespace {
constexpr int NT = VK_NT;
constexpr int MAX_SP = 128;  // absorbers + scatterers staged per layer

struct LayerCoef {
  double chi, xi, phi, g_p, g_m;
};

// Two-stream coefficients of one layer (photo.py:167-203).
__device__ inline LayerCoef two_stream(double w0, double dtau, double mu_ang, double edd,
                                       double ag0, int ag0_zero, double* tran_out) {
  LayerCoef c;
  double tran, zeta_p, zeta_m, ll;
  if (ag0_zero) {
    const double s = sqrt(1.0 - w0);
    tran = exp(-1.0 / edd * s * dtau);
    zeta_p = 0.5 * (1.0 + s);
    zeta_m = 0.5 * (1.0 - s);
    ll = -1.0 * w0 / (1.0 / (mu_ang * mu_ang) - 1.0 / (edd * edd) * (1.0 - w0));
    c.g_p = 0.5 * (ll * (1.0 / edd + 1.0 / mu_ang));
    c.g_m = 0.5 * (ll * (1.0 / edd - 1.0 / mu_ang));
  } else {
    const double wg = 1.0 - w0 * ag0;
    tran = exp(-1.0 / edd * sqrt((1.0 - w0 * ag0) * (1.0 - w0)) * dtau);
    zeta_p = 0.5 * (1.0 + sqrt((1.0 - w0) / wg));
    zeta_m = 0.5 * (1.0 - sqrt((1.0 - w0) / wg));
    ll = ((1.0 - w0) * wg - 1.0) / (1.0 / (mu_ang * mu_ang) - 1.0 / (edd * edd) * (1.0 - w0) * wg);
    c.g_p = 0.5 * (ll * (1.0 / edd + 1.0 / (mu_ang * wg)) + w0 * ag0 * mu_ang / wg);
    c.g_m = 0.5 * (ll * (1.0 / edd - 1.0 / (mu_ang * wg)) - w0 * ag0 * mu_ang / wg);
  }
  c.chi = zeta_m * zeta_m * tran * tran - zeta_p * zeta_p;
  c.xi = zeta_p * zeta_m * (1.0 - tran * tran);
  c.phi = (zeta_m * zeta_m - zeta_p * zeta_p) * tran;
  *tran_out = tran;
  return c;
}

}  // namespace

__global__ __launch_bounds__(NT) void k_photo_optics(const VkCol* __restrict__ cols, int ncols,
                                                     vk_photo_setup P, vk_photo_state S,
                                                     const double* y) {
  __shared__ double dens_a[MAX_SP];
  __shared__ double dens_s[MAX_SP];
  __shared__ double dens_t[MAX_SP];
  const int c = blockIdx.y;
  if (c >= ncols) return;
  if (S.due && !S.due[c]) return;
  if (S.status && (S.status[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  const int nz = col.nz, ni = col.ni, nbin = P.nbin;
  const int b = blockIdx.x * NT + threadIdx.x;
  const bool act = b < nbin;
  const int64_t op = P.off_p[c];
  const double* yc = y + col.off_s;
  const double* dz = P.dz + P.off_l[c];
  double* tau = S.tau + op;
  double* w0a = S.w0 + op;
  const double* crossT = P.absp_T_cross + P.off_l[c] * P.n_absT * nbin;
  // ---- optical depth (top-down cumulative) and single-scattering albedo ----
  double acc = 0.0;
  if (act) tau[(int64_t)nz * nbin + b] = 0.0;
  for (int z = nz - 1; z >= 0; --z) {
    __syncthreads();
    for (int s = threadIdx.x; s < P.n_abs; s += NT) dens_a[s] = yc[(int64_t)z * ni + P.absp_idx[s]];
    for (int s = threadIdx.x; s < P.n_scat; s += NT) dens_s[s] = yc[(int64_t)z * ni + P.scat_idx[s]];
    for (int s = threadIdx.x; s < P.n_absT; s += NT) dens_t[s] = yc[(int64_t)z * ni + P.absp_T_idx[s]];
    __syncthreads();
    if (act) {
      double sa = 0.0, ss = 0.0, st = 0.0;
      for (int s = 0; s < P.n_abs; ++s) sa = fma(dens_a[s], P.absp_cross[(int64_t)s * nbin + b], sa);
      // T-dependent absorbers: cross section of this layer (photo.py:107-111)
      for (int s = 0; s < P.n_absT; ++s)
        st = fma(dens_t[s], crossT[((int64_t)s * nz + z) * nbin + b], st);
      for (int s = 0; s < P.n_scat; ++s) ss = fma(dens_s[s], P.scat_cross[(int64_t)s * nbin + b], ss);
      // photo.py:100-122: ((sum_abs * dz) + (sum_absT * dz)) + (sum_scat * dz),
      // cumulated from the top
      double layer = sa * dz[z];
      if (P.n_absT > 0) layer = layer + st * dz[z];
      layer = layer + ss * dz[z];
      acc = acc + layer;
      tau[(int64_t)z * nbin + b] = acc;
      // photo.py:153-169: w0 from mixing ratios; the ratio is scale-free so
      // the number densities of the layer give it directly
      const double tot_abs = P.n_absT > 0 ? sa + st : sa;
      double w0 = ss / fmax(tot_abs + ss, 1e-300);
      if (isnan(w0)) w0 = 0.0;
      w0 = fmin(w0, 1.0 - 1e-8);
      w0a[(int64_t)z * nbin + b] = w0;
    }
  }
}

__global__ __launch_bounds__(NT) void k_photo_flux(const VkCol* __restrict__ cols, int ncols,
                                                   vk_photo_setup P, vk_photo_state S) {
  const int c = blockIdx.y;
  if (c >= ncols) return;
  if (S.due && !S.due[c]) return;
  if (S.status && (S.status[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  const int nz = col.nz, nbin = P.nbin;
  const int b = blockIdx.x * NT + threadIdx.x;
  const bool act = b < nbin;
  const int64_t op = P.off_p[c];
  const double* tau = S.tau + op;
  const double* w0a = S.w0 + op;
  const double mu_ang = -1.0 * P.mu_zenith;
  double amax = 0.0;
  if (act) {
  // ---- down sweep with the previous call's dflux_u (photo.py:208-223) ------
  double* sfl = S.sflux + op;
  double* dd = S.dflux_d + op;
  double* du = S.dflux_u + op;
  double* af = S.aflux + op;
  const double Ftop = P.sflux_top[b];
  const double edd = P.edd;
  dd[(int64_t)nz * nbin + b] = 0.0;
  double dnext = 0.0;
  double dir_hi = Ftop * exp(tau[(int64_t)nz * nbin + b] / mu_ang) * (-mu_ang);
  sfl[(int64_t)nz * nbin + b] = Ftop * exp(tau[(int64_t)nz * nbin + b] / mu_ang);
  for (int j = nz - 1; j >= 0; --j) {
    const double tj = tau[(int64_t)j * nbin + b];
    const double sj = Ftop * exp(tj / mu_ang);
    sfl[(int64_t)j * nbin + b] = sj;
    const double dir_lo = sj * (-mu_ang);
    double tran;
    const LayerCoef k = two_stream(w0a[(int64_t)j * nbin + b], tj - tau[(int64_t)(j + 1) * nbin + b],
                                   mu_ang, edd, P.ag0, P.ag0_zero, &tran);
    const double i_d = k.phi * k.g_m * dir_hi - (k.chi * k.g_m + k.xi * k.g_p) * dir_lo;
    const double dj = (1.0 / k.chi) * (k.phi * dnext - k.xi * du[(int64_t)j * nbin + b] + i_d / mu_ang);
    dd[(int64_t)j * nbin + b] = dj;
    dnext = dj;
    dir_hi = dir_lo;
  }
  // ---- up sweep (photo.py:225-236) and actinic flux (photo.py:238-244) ------
  double uprev = 0.0;
  du[b] = 0.0;
  const double scale = P.hc / P.bins[b];
  for (int j = 1; j <= nz; ++j) {
    const double tlo = tau[(int64_t)(j - 1) * nbin + b], thi = tau[(int64_t)j * nbin + b];
    const double dir_lo = sfl[(int64_t)(j - 1) * nbin + b] * (-mu_ang);
    const double dir_hi2 = sfl[(int64_t)j * nbin + b] * (-mu_ang);
    double tran;
    const LayerCoef k = two_stream(w0a[(int64_t)(j - 1) * nbin + b], tlo - thi, mu_ang, edd,
                                   P.ag0, P.ag0_zero, &tran);
    const double i_u = k.phi * k.g_p * dir_lo - (k.xi * k.g_m + k.chi * k.g_p) * dir_hi2;
    const double uj = (1.0 / k.chi) * (k.phi * uprev - k.xi * dd[(int64_t)j * nbin + b] + i_u / mu_ang);
    du[(int64_t)j * nbin + b] = uj;
    // layer j-1 actinic flux
    const int l = j - 1;
    const double ave_dir = 0.5 * (sfl[(int64_t)l * nbin + b] + sfl[(int64_t)(l + 1) * nbin + b]);
    const double dif = 0.5 * (uprev + uj + dd[(int64_t)(l + 1) * nbin + b] + dd[(int64_t)l * nbin + b]) / edd;
    const double a_new = (ave_dir + dif) / scale;
    // aflux change vs the previous call (outer_loop.py:330-336)
    const double a_old = af[(int64_t)l * nbin + b];
    if (a_new > P.flux_atol) {
      const double r = fabs(a_new - a_old) / fmax(fabs(a_new), 1e-300);
      amax = r > amax ? r : amax;
    }
    af[(int64_t)l * nbin + b] = a_new;
    uprev = uj;
  }
  }
  // column-wide max of the aflux change (max is order independent)
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(amax, off, 64);
    amax = o > amax ? o : amax;
  }
  if ((threadIdx.x & 63) == 0 && S.aflux_change)
    atomicMax((unsigned long long*)&S.aflux_change[c], (unsigned long long)__double_as_longlong(amax));
}

// J of the non-T branches on FP64 MFMA (photo.py:309-330) and the k rows
// (photo.py:455-478).  v_mfma_f64_16x16x4f64 operands: lane (lr = lane/16,
// lc = lane%16) holds A[row lc][k lr] and B[k lr][col lc]; result item v is
// element (lr + 4v, lc).  Here A = cross_J rows (branches), B = the flux
// weighted by the bin spacing, k = bins, columns = layers.
typedef double ph_d4 __attribute__((ext_vector_type(4)));

// One workgroup per (16-layer tile, column): its four waves take four
// contiguous slabs of the bin chunks for every 16-branch tile of a 64-branch
// group, two chunks in flight per iteration, and the slab sums are added in
// wave order through LDS.
__global__ __launch_bounds__(NT) void k_photo_J(const VkCol* __restrict__ cols, int ncols,
                                                vk_photo_setup P, vk_photo_state S) {
  __shared__ ph_d4 part[NT / 64][4][64];
  const int c = blockIdx.y, t0 = blockIdx.x;
  if (c >= ncols) return;
  if (S.due && !S.due[c]) return;
  if (S.status && (S.status[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  const int nz = col.nz, nbin = P.nbin, d12 = P.din12, n_br = P.n_br;
  if (t0 * 16 >= nz) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, lr = lane >> 4, lc = lane & 15;
  const double* af = S.aflux + P.off_p[c];
  const double* cr = P.cross_J;
  const int nchunk = (nbin + 3) / 4;
  const int ch0 = (int)((int64_t)nchunk * wave / (NT / 64));
  const int ch1 = (int)((int64_t)nchunk * (wave + 1) / (NT / 64));
  const int zrow = min(t0 * 16 + lc, nz - 1);
  const bool okz = t0 * 16 + lc < nz;
  for (int m0 = 0; m0 < n_br; m0 += 64) {
    ph_d4 acc[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) acc[mt] = ph_d4{0.0, 0.0, 0.0, 0.0};
    // this lane's rows: branches m0 + 16 mt + lc (A) and layer t0 * 16 + lc (B)
    int arow[4];
    bool okb[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      arow[mt] = min(m0 + 16 * mt + lc, n_br - 1);
      okb[mt] = m0 + 16 * mt + lc < n_br;
    }
    auto chunk = [&](int ch, double (&av)[4], double& bw) {
      const int bb = 4 * ch + lr;
      const bool inb = bb < nbin;
      const int b = inb ? bb : nbin - 1;
      const double wgt = inb ? (b < d12 ? P.dbin1 : P.dbin2) : 0.0;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) av[mt] = okb[mt] ? cr[(int64_t)arow[mt] * nbin + b] : 0.0;
      bw = okz ? af[(int64_t)zrow * nbin + b] * wgt : 0.0;
    };
    int ch = ch0;
#pragma unroll 1
    for (; ch + 1 < ch1; ch += 2) {
      double a0[4], a1[4], b0, b1;
      chunk(ch, a0, b0);
      chunk(ch + 1, a1, b1);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[mt], b0, acc[mt], 0, 0, 0);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[mt], b1, acc[mt], 0, 0, 0);
    }
    if (ch < ch1) {
      double a0[4], b0;
      chunk(ch, a0, b0);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
        acc[mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[mt], b0, acc[mt], 0, 0, 0);
    }
    __syncthreads();   // the previous group's sums are read
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) part[wave][mt][lane] = acc[mt];
    __syncthreads();
    if (wave == 0) {
      const int z = t0 * 16 + lc;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        ph_d4 tot = part[0][mt][lane];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w) tot = tot + part[w][mt][lane];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int br = m0 + 16 * mt + lr + 4 * v;
          if (z < nz && br < n_br) S.J[P.off_l[c] * n_br + (int64_t)br * nz + z] = tot[v];
        }
      }
    }
  }
}

// The half end-point terms of the two trapezoids (photo.py:319-330) taken
// off the MFMA sums, and the k rows (photo.py:455-478); a thread per
// (branch, layer) of a column.
__global__ __launch_bounds__(NT) void k_photo_J_ends(const VkCol* __restrict__ cols, int ncols,
                                                     vk_photo_setup P, vk_photo_state S,
                                                     double* k) {
  const int c = blockIdx.y;
  if (c >= ncols) return;
  if (S.due && !S.due[c]) return;
  if (S.status && (S.status[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  const int nz = col.nz, nbin = P.nbin, d12 = P.din12;
  const int e = blockIdx.x * NT + threadIdx.x;
  if (e >= P.n_br * nz) return;
  const int br = e / nz, z = e - br * nz;
  const double* afz = S.aflux + P.off_p[c] + (int64_t)z * nbin;
  const double* crb = P.cross_J + (int64_t)br * nbin;
  const double e1 = 0.5 * (afz[0] * crb[0] + afz[d12 - 1] * crb[d12 - 1]) * P.dbin1;
  const double e2 = 0.5 * (afz[d12] * crb[d12] + afz[nbin - 1] * crb[nbin - 1]) * P.dbin2;
  double* Jp = S.J + P.off_l[c] * P.n_br + e;
  const double J = (*Jp - e1) - e2;
  *Jp = J;
  if (k && P.branch_active[br])
    k[col.off_k + (int64_t)P.branch_re_idx[br] * nz + z] = J * P.f_diurnal;
}

// J of the temperature-dependent branches (photo.py:334-347): per (layer,
// column) the split trapezoid against this layer's cross section.
__global__ __launch_bounds__(NT) void k_photo_JT(const VkCol* __restrict__ cols, int ncols,
                                                 vk_photo_setup P, vk_photo_state S, double* k) {
  __shared__ double red[2][NT / 64];
  const int c = blockIdx.y;
  const int z = blockIdx.x;
  if (c >= ncols) return;
  if (S.due && !S.due[c]) return;
  if (S.status && (S.status[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  const int nz = col.nz;
  if (z >= nz) return;
  const int nbin = P.nbin, d12 = P.din12, tid = threadIdx.x;
  const double* af = S.aflux + P.off_p[c] + (int64_t)z * nbin;
  for (int br = 0; br < P.n_brT; ++br) {
    const double* cr = P.cross_J_T + (P.off_l[c] * P.n_brT + (int64_t)br * nz + z) * nbin;
    double s1 = 0.0, s2 = 0.0;
    for (int b = tid; b < nbin; b += NT) {
      const double v = af[b] * cr[b];
      if (b < d12) s1 += v; else s2 += v;
    }
    for (int off = 32; off > 0; off >>= 1) {
      s1 += __shfl_xor(s1, off, 64);
      s2 += __shfl_xor(s2, off, 64);
    }
    __syncthreads();
    if ((tid & 63) == 0) { red[0][tid >> 6] = s1; red[1][tid >> 6] = s2; }
    __syncthreads();
    if (tid == 0) {
      const double d1 = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
      const double d2 = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
      const double J1 = d1 * P.dbin1 - 0.5 * (af[0] * cr[0] + af[d12 - 1] * cr[d12 - 1]) * P.dbin1;
      const double J2 = d2 * P.dbin2 - 0.5 * (af[d12] * cr[d12] + af[nbin - 1] * cr[nbin - 1]) * P.dbin2;
      const double J = J1 + J2;
      S.J_T[P.off_l[c] * P.n_brT + (int64_t)br * nz + z] = J;
      if (k && P.branch_T_active[br])
        k[col.off_k + (int64_t)P.branch_T_re_idx[br] * nz + z] = J * P.f_diurnal;
    }
  }
}

__global__ void k_photo_reset_change(int ncols, vk_photo_state S) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols || !S.aflux_change) return;
  if (S.due && !S.due[c]) return;
  S.aflux_change[c] = 0.0;
}

int vk_launch_photo(int parts, const VkCol* cols, int ncols, int max_nz,
                    const vk_photo_setup& P, const vk_photo_state& S, const double* y, double* k,
                    hipStream_t st) {
  if (P.n_abs > MAX_SP || P.n_scat > MAX_SP || P.n_absT > MAX_SP) {
    vk_set_error("photolysis: at most %d absorbers, %d T-dependent absorbers and %d scatterers",
                 MAX_SP, MAX_SP, MAX_SP);
    return VK_ERR_UNSUPPORTED;
  }
  if (P.n_absT < 0 || P.n_brT < 0 || (P.n_absT > 0 && (!P.absp_T_cross || !P.absp_T_idx)) ||
      (P.n_brT > 0 && (!P.cross_J_T || !P.branch_T_re_idx || !P.branch_T_active || !S.J_T))) {
    vk_set_error("photolysis: inconsistent temperature-dependent cross sections");
    return VK_ERR_ARG;
  }
  if (P.din12 < 1 || P.din12 >= P.nbin) {
    vk_set_error("photolysis: din12_indx=%d must lie in 1..nbin-1 (nbin=%d)", P.din12, P.nbin);
    return VK_ERR_ARG;
  }
  const dim3 gb((P.nbin + NT - 1) / NT, ncols);
  if (parts & 1) {
    hipLaunchKernelGGL(k_photo_optics, gb, dim3(NT), 0, st, cols, ncols, P, S, y);
    VK_CHECK_HIP(hipGetLastError());
  }
  if (parts & 2) {
    hipLaunchKernelGGL(k_photo_reset_change, dim3((ncols + 255) / 256), dim3(256), 0, st, ncols, S);
    VK_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_photo_flux, gb, dim3(NT), 0, st, cols, ncols, P, S);
    VK_CHECK_HIP(hipGetLastError());
  }
  if (parts & 4) {
    if (P.n_br > 0) {
      hipLaunchKernelGGL(k_photo_J, dim3((max_nz + 15) / 16, ncols), dim3(NT), 0, st, cols,
                         ncols, P, S);
      VK_CHECK_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_photo_J_ends, dim3((P.n_br * max_nz + NT - 1) / NT, ncols), dim3(NT), 0,
                         st, cols, ncols, P, S, k);
      VK_CHECK_HIP(hipGetLastError());
    }
    if (P.n_brT > 0) {
      hipLaunchKernelGGL(k_photo_JT, dim3(max_nz, ncols), dim3(NT), 0, st, cols, ncols, P, S, k);
      VK_CHECK_HIP(hipGetLastError());
    }
  }
  return VK_OK;
}
