// Rate construction k(T, M) on the device: forward rates, low-T caps,
// NASA-9 Gibbs energies, equilibrium constants and the reverse fill.
//
// Restates rates.compute_forward_k (rates.py:61-122), rates.apply_lowT_caps
// (rates.py:138-167), gibbs.gibbs_sp_vector (gibbs.py:41-78),
// gibbs.K_eq_array (gibbs.py:81-126), gibbs.fill_reverse_k
// (gibbs.py:129-152) and rates.build_rate_array (rates.py:185-207).  The
// host classifies each row once (vk_rates_create): the forward form
// (modified Arrhenius, Lindemann falloff, the hard-coded Troe reaction),
// the low-T cap it takes, and whether its reverse is filled from K_eq.
// Every output entry (row or species, layer) is one thread; a reverse row
// recomputes its forward rate and K_eq itself, so a full build is one
// launch with no intermediate arrays.
#include "vk_common.h"

namespace {

__device__ inline double arrh(double a, double n, double E, double T) {
  return a * pow(T, n) * exp(-E / T);
}

// OH + CH3 + M -> CH3OH + M (rates.py:41-58).
__device__ inline double troe_oh_ch3(double T, double M) {
#pragma clang fp contract(off)
  const double k0 = 1.932e3 * pow(T, -9.88) * exp(-7544.0 / T) +
                    5.109e-11 * pow(T, -6.25) * exp(-1433.0 / T);
  const double kinf = 1.031e-10 * pow(T, -0.018) * exp(16.74 / T);
  const double Fc =
      0.1855 * exp(-T / 155.8) + 0.8145 * exp(-T / 1675.0) + exp(-4531.0 / T);
  const double nn = 0.75 - 1.27 * log(Fc);
  const double q = log(k0 * M / kinf) / nn;
  const double ff = exp(log(Fc) / (1.0 + q * q));
  return k0 / (1.0 + k0 * M / kinf) * ff;
}

// Forward rate of row i (0 for rows the reference leaves at zero), with the
// low-T cap applied when requested (rates.py:108-120, 149-165).
__device__ inline double forward_rate(const VkRatesDev& R, int i, double T, double M,
                                      int lowt_on) {
#pragma clang fp contract(off)
  const double* p = R.par + 6 * i;
  double k = 0.0;
  switch (R.form[i]) {
    case VK_KF_ARRHENIUS:
      k = arrh(p[0], p[1], p[2], T);
      break;
    case VK_KF_LINDEMANN: {
      const double k0 = arrh(p[0], p[1], p[2], T);
      const double kinf = arrh(p[3], p[4], p[5], T);
      k = k0 / (1.0 + k0 * M / kinf);
    } break;
    case VK_KF_TROE_OH_CH3:
      k = troe_oh_ch3(T, M);
      break;
    default:
      break;
  }
  return lowt_on ? vk_lowt_cap(R.lowt[i], k, T, M) : k;
}

// g/(RT) of species j from its NASA-9 fit: H/RT - S/R, the low range below
// 1000 K (gibbs.py:47-76).
__device__ inline double gibbs_rt(const double* c, double T) {
#pragma clang fp contract(off)
  const double* a = c + (T < 1000.0 ? 0 : 10);
  const double T2 = T * T, T3 = pow(T, 3.0), T4 = pow(T, 4.0), lT = log(T);
  const double h = -a[0] / T2 + a[1] * lT / T + a[2] + a[3] * T / 2.0 + a[4] * T2 / 3.0 +
                   a[5] * T3 / 4.0 + a[6] * T4 / 5.0 + a[8] / T;
  const double s = -a[0] / T2 / 2.0 - a[1] / T + a[2] * lT + a[3] * T + a[4] * T2 / 2.0 +
                   a[5] * T3 / 3.0 + a[6] * T4 / 4.0 + a[9];
  return h - s;
}

// Equilibrium constant of forward row i; g(j) gives g/RT of species j, 0
// for the padding species (gibbs.py:100-124).
template <class G>
__device__ inline double keq(const VkRatesDev& R, int i, double T, G g) {
#pragma clang fp contract(off)
  if (!R.has_rev[i]) return 1.0;
  double rs = 0.0, ps = 0.0, nr = 0.0, np = 0.0;
  for (int s = 0; s < R.T; ++s) {
    const double st = R.r_st[i * R.T + s];
    if (st != 0.0) rs = rs + st * g(R.r_idx[i * R.T + s]);
  }
  for (int s = 0; s < R.T; ++s) {
    const double st = R.p_st[i * R.T + s];
    if (st != 0.0) ps = ps + st * g(R.p_idx[i * R.T + s]);
  }
  for (int s = 0; s < R.T; ++s) nr = nr + R.r_st[i * R.T + s];
  for (int s = 0; s < R.T; ++s) np = np + R.p_st[i * R.T + s];
  const double dn = nr - np;
  double K = exp(rs - ps);
  if (dn != 0.0) K = K * pow(VK_GIBBS_CORR * T, dn);
  return K;
}

// Row i of the full build at one layer (rates.py:185-207): forward rows,
// reverse rows k_f / K_eq below stop_rev_indx, the rows past it zeroed
// (gibbs.py:149-150), then the host's zero mask.
__device__ inline double build_entry(const VkRatesDev& R, int i, double T, double M,
                                     int lowt_on) {
#pragma clang fp contract(off)
  if (R.zero_rows && R.zero_rows[i]) return 0.0;
  if (i == 0) return 0.0;
  if (i >= R.stop_rev + 1 && ((i - R.stop_rev - 1) & 1) == 0) return 0.0;
  if ((i & 1) == 0) {
    if (i >= R.stop_rev) return 0.0;
    const double f = forward_rate(R, i - 1, T, M, lowt_on);
    const double K = keq(R, i - 1, T, [&](int j) {
      return j < R.ni ? gibbs_rt(R.nasa9 + 20 * j, T) : 0.0;
    });
    return K > 0.0 ? f / K : 0.0;
  }
  return forward_rate(R, i, T, M, lowt_on);
}

// One column: out (rows or species, nz), entry (i, z) at i * nz + z.
__global__ __launch_bounds__(256) void k_rates_eval(VkRatesDev R, int what, int nz,
                                                    const double* __restrict__ Tz,
                                                    const double* __restrict__ Mz, int lowt_on,
                                                    const double* __restrict__ in_a,
                                                    const double* __restrict__ in_b,
                                                    double* __restrict__ out) {
#pragma clang fp contract(off)
  const int rows = (what == VK_RATES_GIBBS) ? R.ni : R.nr + 1;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)rows * nz) return;
  const int i = (int)(e / nz), z = (int)(e % nz);
  const double T = Tz[z];
  const double M = Mz ? Mz[z] : 0.0;
  double v = 0.0;
  switch (what) {
    case VK_RATES_FORWARD:
      v = (i & 1) ? forward_rate(R, i, T, M, lowt_on) : 0.0;
      break;
    case VK_RATES_CAPS:
      v = (i & 1) ? vk_lowt_cap(R.lowt[i], in_a[e], T, M) : in_a[e];
      break;
    case VK_RATES_GIBBS:
      v = gibbs_rt(R.nasa9 + 20 * i, T);
      break;
    case VK_RATES_KEQ:
      v = (i & 1) ? keq(R, i, T, [&](int j) { return j < R.ni ? in_a[(int64_t)j * nz + z] : 0.0; })
                  : 1.0;
      break;
    case VK_RATES_FILL: {
      v = in_a[e];
      if (R.zero_rows && R.zero_rows[i]) {
        v = 0.0;
      } else if (i >= 2 && i < R.stop_rev && (i & 1) == 0) {
        const double K = in_b[e - nz];
        v = K > 0.0 ? in_a[e - nz] / K : 0.0;
      } else if (i >= R.stop_rev + 1 && ((i - R.stop_rev - 1) & 1) == 0) {
        v = 0.0;
      }
    } break;
    default:  // VK_RATES_BUILD
      v = build_entry(R, i, T, M, lowt_on);
      break;
  }
  out[e] = v;
}

// A batch: column c's rates (its network's S rows x nz layers) at off_k,
// T and M read from its packed atmosphere record.
__global__ __launch_bounds__(256) void k_rates_build_batch(const VkCol* __restrict__ cols,
                                                           int ncols,
                                                           const VkRatesDev* __restrict__ rates,
                                                           const double* __restrict__ atm,
                                                           int lowt_on, double* __restrict__ k) {
  const int c = blockIdx.y;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  const VkRatesDev& R = rates[col.net];
  const int nz = col.nz;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)(R.nr + 1) * nz) return;
  const int i = (int)(e / nz), z = (int)(e % nz);
  const VkAtmOff o = vk_atm_offsets(nz, col.ni);
  const double* rec = atm + col.off_a;
  k[col.off_k + e] = build_entry(R, i, rec[o.Tco + z], rec[o.M + z], lowt_on);
}

}  // namespace

int vk_launch_rates_eval(const VkRatesDev& R, int what, int nz, const double* T, const double* M,
                         int lowt_on, const double* in_a, const double* in_b, double* out,
                         hipStream_t stream) {
  const int rows = (what == VK_RATES_GIBBS) ? R.ni : R.nr + 1;
  const int64_t n = (int64_t)rows * nz;
  if (n <= 0) return VK_OK;
  hipLaunchKernelGGL(k_rates_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, R,
                     what, nz, T, M, lowt_on, in_a, in_b, out);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_rates_build_batch(const VkCol* cols, int ncols, int max_S, int max_nz,
                                const VkRatesDev* rates, const double* atm, int lowt_on,
                                double* k, hipStream_t stream) {
  const int64_t n = (int64_t)max_S * max_nz;
  if (n <= 0 || ncols <= 0) return VK_OK;
  hipLaunchKernelGGL(k_rates_build_batch, dim3((unsigned)((n + 255) / 256), ncols), dim3(256), 0,
                     stream, cols, ncols, rates, atm, lowt_on, k);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
