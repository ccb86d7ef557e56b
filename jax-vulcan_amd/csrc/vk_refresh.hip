// Atmosphere refresh on the device: mean molecular mass, the hydrostatic
// column (g, Hp, dz, zco) and the interface averages from the current
// mixing ratios, and the diffusion-limited escape flux at the top.
//
// Restates atm_refresh.update_mu_dz_jax (atm_refresh.py:36-98) and
// update_phi_esc_jax (atm_refresh.py:101-122).  One 256-thread workgroup
// per column: the mean masses are layer-parallel; the column itself is a
// true prefix dependency (z_{i+1} = z_i + dz_i with g depending on z_i) and
// runs on one lane, up from the reference layer and then down from it,
// O(nz) steps once every update_frq accepted steps; the interface averages
// and the escape fluxes are parallel again.
#include "vk_common.h"

namespace {

constexpr int NT = 256;

}  // namespace

// ymix / y as given, or (sol non-null) the controller's clip of the step
// output sol against the current ymix: ymix_new = clip / gas column density
// and the clipped top-layer densities.
__global__ __launch_bounds__(NT) void k_atm_refresh(const VkCol* __restrict__ cols, int ncols,
                                                    int mode, const double* __restrict__ ymix_in,
                                                    const double* __restrict__ y_in,
                                                    const double* __restrict__ sol_in,
                                                    vk_ss_params p, double* atm,
                                                    vk_refresh_state r) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (r.due && r.due[c] == 0) return;
  const VkCol col = cols[c];
  const int nz = col.nz, ni = col.ni, tid = threadIdx.x;
  const VkAtmOff ao = vk_atm_offsets(nz, ni);
  const VkRefOff ro = vk_refresh_offsets(nz, ni);
  const double* rec = r.rec + r.off_r[c];
  const double* pico = rec + ro.pico;
  const double* mass = rec + ro.mass;
  const double* esc = rec + ro.esc;
  const int pref = (int)rec[ro.scal + 0];
  if (pref < 0 || pref >= nz || nz < 2) return;  // vk_atm_refresh_f64 checks records it packs
  const double zpref = rec[ro.scal + 1], gs = rec[ro.scal + 2], Rp = rec[ro.scal + 3];
  const double max_flux = rec[ro.scal + 4], kb = rec[ro.scal + 5], Navo = rec[ro.scal + 6];
  double* a = atm + col.off_a;
  const double* Tco = a + ao.Tco;
  double* geo = r.geo + r.off_g[c];
  double* g_mu = geo;
  double* g_g = geo + nz;
  double* g_Hp = geo + 2 * nz;
  double* g_dz = geo + 3 * nz;
  double* g_zco = geo + 4 * nz;
  double* mu = lds;              // (nz)
  double* g = mu + nz;
  double* Hp = g + nz;
  double* dz = Hp + nz;
  const double* gas = a + ao.gas;
  const double* ymix = ymix_in ? ymix_in + col.off_s : nullptr;
  const double* y = y_in ? y_in + col.off_s : nullptr;
  const double* sol = sol_in ? sol_in + col.off_s : nullptr;
  auto clipped = [&](int64_t e) { return vk_clip_val(sol[e], ymix[e], p); };

  if (mode & 1) {
    for (int z = tid; z < nz; z += NT) {
      const int64_t row = (int64_t)z * ni;
      double m = 0.0;
      if (sol) {
        double ys = 0.0;
        for (int i = 0; i < ni; ++i) {
          const double sc = clipped(row + i);
          if (gas[i] != 0.0) ys += sc;
        }
        for (int i = 0; i < ni; ++i) m = m + (clipped(row + i) / ys) * mass[i];
      } else {
        for (int i = 0; i < ni; ++i) m = m + ymix[row + i] * mass[i];
      }
      mu[z] = m;
    }
    __syncthreads();
    if (tid == 0) {
      double zc = zpref;
      g_zco[pref] = zpref;
      for (int i = pref; i < nz; ++i) {   // upward (atm_refresh.py:56-64)
        const double q = Rp / (Rp + zc);
        const double gi = (i == pref) ? gs : gs * (q * q);
        const double hp = kb * Tco[i] / (mu[i] / Navo * gi);
        const double d = hp * log(pico[i] / pico[i + 1]);
        zc = zc + d;
        g[i] = gi;
        Hp[i] = hp;
        dz[i] = d;
        g_zco[i + 1] = zc;
      }
      zc = zpref;
      for (int i = pref - 1; i >= 0; --i) {   // downward (atm_refresh.py:72-78)
        const double q = Rp / (Rp + zc);
        const double gi = gs * (q * q);
        const double hp = kb * Tco[i] / (mu[i] / Navo * gi);
        const double d = hp * log(pico[i] / pico[i + 1]);
        zc = zc - d;
        g[i] = gi;
        Hp[i] = hp;
        dz[i] = d;
        g_zco[i] = zc;
      }
    }
    __syncthreads();
    for (int z = tid; z < nz; z += NT) {
      g_mu[z] = mu[z];
      g_g[z] = g[z];
      g_Hp[z] = Hp[z];
      g_dz[z] = dz[z];
      a[ao.g + z] = g[z];
      if (z + 1 < nz) {
        a[ao.dzi + z] = 0.5 * (dz[z] + dz[z + 1]);
        a[ao.Hpi + z] = 0.5 * (Hp[z] + Hp[z + 1]);
      }
    }
  }
  if (mode & 2) {
    const double g_top = (mode & 1) ? g[nz - 1] : g_g[nz - 1];
    const double hp_top = (mode & 1) ? Hp[nz - 1] : g_Hp[nz - 1];
    const double T_top = Tco[nz - 1];
    const double* Dzz_top = a + ao.Dzz + (int64_t)(nz - 2) * ni;
    const double* ms = a + ao.ms;
    const int64_t top = (int64_t)(nz - 1) * ni;
    for (int i = tid; i < ni; i += NT) {
      if (esc[i] == 0.0) continue;
      const double yt = sol ? clipped(top + i) : y[top + i];
      double f = -Dzz_top[i] * yt * (1.0 / hp_top - ms[i] * g_top / (Navo * kb * T_top));
      f = fmax(f, -max_flux);
      a[ao.top_flux + i] = f;
    }
  }
  if (r.due) {
    __syncthreads();
    if (tid == 0) r.due[c] = 0;
  }
}

int vk_launch_atm_refresh(const VkCol* cols, int ncols, int max_nz, int mode, const double* ymix,
                          const double* y, const double* sol, const vk_ss_params& p, double* atm,
                          const vk_refresh_state& r, hipStream_t stream) {
  if (ncols <= 0) return VK_OK;
  const size_t shm = sizeof(double) * 4 * (size_t)max_nz;
  hipLaunchKernelGGL(k_atm_refresh, dim3(ncols), dim3(NT), shm, stream, cols, ncols, mode, ymix,
                     y, sol, p, atm, r);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
