// Rate construction k(T, M) on the device: forward rates, low-T caps,
// NASA-9 Gibbs energies, equilibrium constants and the reverse fill.
//
// Restates rates.compute_forward_k (rates.py:61-122), rates.apply_lowT_caps
// (rates.py:142-167), gibbs.gibbs_sp_vector (gibbs.py:41-74),
// gibbs.K_eq_array (gibbs.py:77-128), gibbs.fill_reverse_k
// (gibbs.py:131-154) and rates.build_rate_array (rates.py:185-208).  The
// host classifies each (forward, reverse) row pair once into a 128-byte
// record (vk_rates_create): the forward form (modified Arrhenius,
// Lindemann falloff, the hard-coded Troe reaction), the low-T cap it
// takes, whether its reverse is filled from K_eq, which of its rows the
// full build holds at zero, and the stoichiometry K_eq reads.
//
// The separate stages (forward, caps, Gibbs, K_eq, fill) are one thread per
// output entry.  The full build is one launch with no intermediate arrays
// in HBM: a block stages the Gibbs energies of all species over a tile of
// 64 layers in LDS; each wave then takes one row pair at a time, its 64
// lanes on the 64 layers, so the pair's record is a wave-uniform fetch and
// each row's 64 stores are contiguous; the forward rate is evaluated once
// for both rows of the pair.
#include "vk_common.h"

namespace {

// kb / P0, the concentration-unit correction of K_eq (gibbs.py:12-13,
// phy_const.py:3).
constexpr double VK_GIBBS_CORR = 1.38064852e-16 / 1.0e6;

// ---------------------------------------------------------------------------
// Values and dual numbers.  Every rate expression below is written once,
// over a value type V: double for the build, Dual (value, d/dT, d/dM) for
// the vector-Jacobian product.  The double instantiation rounds exactly as
// the plain expression does (fp contraction off everywhere).
// ---------------------------------------------------------------------------

struct Dual {
  double v, t, m;
  Dual() = default;
  __device__ Dual(double x) : v(x), t(0.0), m(0.0) {}
  __device__ Dual(double x, double dt, double dm) : v(x), t(dt), m(dm) {}
};

__device__ inline Dual operator+(Dual a, Dual b) {
#pragma clang fp contract(off)
  return {a.v + b.v, a.t + b.t, a.m + b.m};
}
__device__ inline Dual operator-(Dual a, Dual b) {
#pragma clang fp contract(off)
  return {a.v - b.v, a.t - b.t, a.m - b.m};
}
__device__ inline Dual operator-(Dual a) { return {-a.v, -a.t, -a.m}; }
__device__ inline Dual operator*(Dual a, Dual b) {
#pragma clang fp contract(off)
  return {a.v * b.v, a.t * b.v + a.v * b.t, a.m * b.v + a.v * b.m};
}
__device__ inline Dual operator/(Dual a, Dual b) {
#pragma clang fp contract(off)
  const double q = a.v / b.v;
  return {q, (a.t - q * b.t) / b.v, (a.m - q * b.m) / b.v};
}
__device__ inline Dual operator+(Dual a, double b) { return a + Dual(b); }
__device__ inline Dual operator+(double a, Dual b) { return Dual(a) + b; }
__device__ inline Dual operator-(Dual a, double b) { return a - Dual(b); }
__device__ inline Dual operator-(double a, Dual b) { return Dual(a) - b; }
__device__ inline Dual operator*(Dual a, double b) { return a * Dual(b); }
__device__ inline Dual operator*(double a, Dual b) { return Dual(a) * b; }
__device__ inline Dual operator/(Dual a, double b) { return a / Dual(b); }
__device__ inline Dual operator/(double a, Dual b) { return Dual(a) / b; }

__device__ inline double val(double x) { return x; }
__device__ inline double val(const Dual& x) { return x.v; }
__device__ inline double vexp(double x) { return exp(x); }
__device__ inline double vlog(double x) { return log(x); }
__device__ inline double vpow(double x, double c) { return pow(x, c); }
__device__ inline Dual vexp(Dual x) {
#pragma clang fp contract(off)
  const double e = exp(x.v);
  return {e, e * x.t, e * x.m};
}
__device__ inline Dual vlog(Dual x) {
#pragma clang fp contract(off)
  return {log(x.v), x.t / x.v, x.m / x.v};
}
__device__ inline Dual vpow(Dual x, double c) {
#pragma clang fp contract(off)
  const double p = pow(x.v, c);
  const double d = c * pow(x.v, c - 1.0);
  return {p, d * x.t, d * x.m};
}

// ---------------------------------------------------------------------------
// Rate expressions
// ---------------------------------------------------------------------------

// a T^n exp(-E/T) as a exp(n ln T - E/T) with ln T shared by the layer: one
// exponential instead of pow and exp, within 3e-14 relative of the
// two-function form for |E/T| <= 700 (rates.py:36-38).
template <class V>
__device__ inline V arrh(double a, double n, double E, V T, V lnT) {
#pragma clang fp contract(off)
  return a * vexp(n * lnT - E / T);
}

// OH + CH3 + M -> CH3OH + M (rates.py:41-58).
template <class V>
__device__ inline V troe_oh_ch3(V T, V M) {
#pragma clang fp contract(off)
  const V k0 = 1.932e3 * vpow(T, -9.88) * vexp(-7544.0 / T) +
               5.109e-11 * vpow(T, -6.25) * vexp(-1433.0 / T);
  const V kinf = 1.031e-10 * vpow(T, -0.018) * vexp(16.74 / T);
  const V Fc = 0.1855 * vexp(-T / 155.8) + 0.8145 * vexp(-T / 1675.0) + vexp(-4531.0 / T);
  const V nn = 0.75 - 1.27 * vlog(Fc);
  const V q = vlog(k0 * M / kinf) / nn;
  const V ff = vexp(vlog(Fc) / (1.0 + q * q));
  return k0 / (1.0 + k0 * M / kinf) * ff;
}

// The low-T caps of rates.apply_lowT_caps (rates.py:149-165).
template <class V>
__device__ inline V lowt_cap(int code, V k, V T, V M) {
#pragma clang fp contract(off)
  switch (code) {
    case VK_LOWT_CH3:
      if (val(T) <= 277.5) {
        const double k0 = 6.0e-29;
        const V kinf = 2.06e-10 * vpow(T, -0.4);
        return k0 / (1.0 + k0 * M / kinf);
      }
      return k;
    case VK_LOWT_C2H4:
      return val(T) <= 300.0 ? V(3.7e-30) : k;
    case VK_LOWT_C2H5:
      return val(T) <= 200.0 ? V(2.49e-27) : k;
    default:
      return k;
  }
}

// Forward rate of a pair's forward row (0 for rows the reference leaves at
// zero), with the low-T cap applied when requested (rates.py:108-120,
// 149-165).
template <class V>
__device__ inline V forward_rate(const VkRatePair& q, V T, V lnT, V M, int lowt_on) {
#pragma clang fp contract(off)
  V k = V(0.0);
  switch (q.form) {
    case VK_KF_ARRHENIUS:
      k = arrh(q.par[0], q.par[1], q.par[2], T, lnT);
      break;
    case VK_KF_LINDEMANN: {
      const V k0 = arrh(q.par[0], q.par[1], q.par[2], T, lnT);
      const V kinf = arrh(q.par[3], q.par[4], q.par[5], T, lnT);
      k = k0 / (1.0 + k0 * M / kinf);
    } break;
    case VK_KF_TROE_OH_CH3:
      k = troe_oh_ch3(T, M);
      break;
    default:
      break;
  }
  return lowt_on ? lowt_cap(q.lowt, k, T, M) : k;
}

// Powers of T shared by every species' Gibbs fit at one layer.
template <class V>
struct TPow {
  V T, T2, T3, T4, lT;
};

template <class V>
__device__ inline TPow<V> tpow(V T) {
#pragma clang fp contract(off)
  return TPow<V>{T, T * T, vpow(T, 3.0), vpow(T, 4.0), vlog(T)};
}

// g/(RT) of species j from its NASA-9 fit: H/RT - S/R, the low range below
// 1000 K (gibbs.py:49-74).
template <class V>
__device__ inline V gibbs_rt(const double* c, const TPow<V>& p) {
#pragma clang fp contract(off)
  const V T = p.T, T2 = p.T2, T3 = p.T3, T4 = p.T4, lT = p.lT;
  const double* a = c + (val(T) < 1000.0 ? 0 : 10);
  const V h = -a[0] / T2 + a[1] * lT / T + a[2] + a[3] * T / 2.0 + a[4] * T2 / 3.0 +
              a[5] * T3 / 4.0 + a[6] * T4 / 5.0 + a[8] / T;
  const V s = -a[0] / T2 / 2.0 - a[1] / T + a[2] * lT + a[3] * T + a[4] * T2 / 2.0 +
              a[5] * T3 / 3.0 + a[6] * T4 / 4.0 + a[9];
  return h - s;
}

// Equilibrium constant of a pair's forward row; g(j) gives g/RT of species
// j (the padding species counts 0); lct = ln(CORR T) of the layer
// (gibbs.py:107-126).
template <class V, class G>
__device__ inline V keq(const VkRatePair& q, V T, V lct, int ni, G g) {
#pragma clang fp contract(off)
  if (!(q.flags & VK_PAIR_HAS_REV)) return V(1.0);
  V rs = V(0.0), ps = V(0.0);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double st = q.st[s];
    if (st != 0.0) rs = rs + st * (q.sp[s] < ni ? g(q.sp[s]) : V(0.0));
  }
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double st = q.st[4 + s];
    if (st != 0.0) ps = ps + st * (q.sp[4 + s] < ni ? g(q.sp[4 + s]) : V(0.0));
  }
  V K = vexp(rs - ps);
  // (CORR T)^dn: exact for dn = +-1, exp(dn ln(CORR T)) otherwise.
  if (q.dn == 1.0) {
    K = K * (VK_GIBBS_CORR * T);
  } else if (q.dn == -1.0) {
    K = K * (1.0 / (VK_GIBBS_CORR * T));
  } else if (q.dn != 0.0) {
    K = K * vexp(q.dn * lct);
  }
  return K;
}

// Both rows of a pair in the full build at one layer (rates.py:185-208):
// the forward rate, and k_f / K_eq for a reverse row below stop_rev_indx
// (0 where K <= 0, gibbs.py:141-149); the rows the record marks are zero.
template <class V, class G>
__device__ inline void build_pair(const VkRatePair& q, V T, V lnT, V lct, V M, int lowt_on,
                                  int ni, G g, V& k_fwd, V& k_rev) {
#pragma clang fp contract(off)
  const bool zf = q.flags & VK_PAIR_ZERO_FWD;
  const bool zr = q.flags & VK_PAIR_ZERO_REV;
  V f = V(0.0);
  if (!zf || !zr) f = forward_rate(q, T, lnT, M, lowt_on);
  k_fwd = zf ? V(0.0) : f;
  k_rev = V(0.0);
  if (!zr) {
    const V K = keq(q, T, lct, ni, g);
    k_rev = val(K) > 0.0 ? f / K : V(0.0);
  }
}

// One column, one stage: out (rows or species, nz), entry (i, z) at
// i * nz + z.
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
  const bool fwd_row = (what != VK_RATES_GIBBS) && (i & 1);
  double v = 0.0;
  switch (what) {
    case VK_RATES_FORWARD:
      v = fwd_row ? forward_rate(R.pairs[i >> 1], T, log(T), M, lowt_on) : 0.0;
      break;
    case VK_RATES_CAPS:
      v = fwd_row ? lowt_cap(R.pairs[i >> 1].lowt, in_a[e], T, M) : in_a[e];
      break;
    case VK_RATES_GIBBS:
      v = gibbs_rt(R.nasa9 + 20 * i, tpow(T));
      break;
    case VK_RATES_KEQ:
      v = fwd_row ? keq(R.pairs[i >> 1], T, log(VK_GIBBS_CORR * T), R.ni,
                        [&](int j) { return in_a[(int64_t)j * nz + z]; })
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
    default:
      break;
  }
  out[e] = v;
}

constexpr int VK_RATES_ZT = 64;   // layers per block, one per lane
constexpr int VK_RATES_NT = 512;  // threads per block

// The full build over layers z0 .. z0+63 of one column.
__device__ inline void build_tile(const VkRatesDev& R, const double* __restrict__ Tz,
                                  const double* __restrict__ Mz, int nz, int z0, int lowt_on,
                                  double* __restrict__ out, double* lds) {
#pragma clang fp contract(off)
  double* lT = lds;                       // per-layer T, powers, logs, M
  double* lT2 = lT + VK_RATES_ZT;
  double* lT3 = lT2 + VK_RATES_ZT;
  double* lT4 = lT3 + VK_RATES_ZT;
  double* llog = lT4 + VK_RATES_ZT;
  double* lM = llog + VK_RATES_ZT;
  double* lct = lM + VK_RATES_ZT;         // ln(CORR T)
  double* gl = lct + VK_RATES_ZT;         // (ni, ZT) g/RT
  const int tid = threadIdx.x;
  const int zn = min(VK_RATES_ZT, nz - z0);
  if (tid < VK_RATES_ZT) {
    const double T = tid < zn ? Tz[z0 + tid] : 300.0;
    const TPow<double> p = tpow(T);
    lT[tid] = p.T;
    lT2[tid] = p.T2;
    lT3[tid] = p.T3;
    lT4[tid] = p.T4;
    llog[tid] = p.lT;
    lM[tid] = tid < zn ? Mz[z0 + tid] : 0.0;
    lct[tid] = log(VK_GIBBS_CORR * T);
  }
  __syncthreads();
  for (int e = tid; e < R.ni * VK_RATES_ZT; e += VK_RATES_NT) {
    const int j = e / VK_RATES_ZT, z = e % VK_RATES_ZT;
    const TPow<double> p{lT[z], lT2[z], lT3[z], lT4[z], llog[z]};
    gl[e] = gibbs_rt(R.nasa9 + 20 * j, p);
  }
  __syncthreads();
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  if (lane >= zn) return;
  const double T = lT[lane], M = lM[lane], lnT = llog[lane], lc = lct[lane];
  auto g = [&](int j) { return gl[j * VK_RATES_ZT + lane]; };
  double* o = out + z0 + lane;
  if (wave == 0) o[0] = 0.0;               // row 0
  // The records are read-only for the kernel's lifetime and the index is
  // wave-uniform: read through the constant address space they arrive in
  // scalar registers.
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  using cline_t = const __attribute__((address_space(4))) u32x4*;
  const cline_t lines = (cline_t)R.pairs;
  for (int p = wave; p < R.npairs; p += VK_RATES_NT / 64) {
    VkRatePair q;
    u32x4* qv = reinterpret_cast<u32x4*>(&q);
#pragma unroll
    for (int k = 0; k < 8; ++k) qv[k] = lines[8 * p + k];
    double kf, kr;
    build_pair(q, T, lnT, lc, M, lowt_on, R.ni, g, kf, kr);
    const int i = 2 * p + 1;
    o[(int64_t)i * nz] = kf;
    if (i + 1 <= R.nr) o[(int64_t)(i + 1) * nz] = kr;
  }
}

// One column, T and M given as arrays; out (nr+1, nz).
__global__ __launch_bounds__(VK_RATES_NT) void k_rates_build_one(
    VkRatesDev R, int nz, const double* __restrict__ Tz, const double* __restrict__ Mz,
    int lowt_on, double* __restrict__ out) {
  extern __shared__ double lds[];
  build_tile(R, Tz, Mz, nz, blockIdx.x * VK_RATES_ZT, lowt_on, out, lds);
}

// A batch: column c's rates (its network's S rows x nz layers) at off_k,
// T and M read from its packed atmosphere record.
__global__ __launch_bounds__(VK_RATES_NT) void k_rates_build_batch(
    const VkCol* __restrict__ cols, int ncols, const VkRatesDev* __restrict__ rates,
    const double* __restrict__ atm, int lowt_on, double* __restrict__ k) {
  extern __shared__ double lds[];
  const int c = blockIdx.y;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  const int z0 = blockIdx.x * VK_RATES_ZT;
  if (z0 >= col.nz) return;
  const VkAtmOff o = vk_atm_offsets(col.nz, col.ni);
  const double* rec = atm + col.off_a;
  const VkRatesDev R = rates[col.net];
  build_tile(R, rec + o.Tco, rec + o.M, col.nz, z0, lowt_on, k + col.off_k, lds);
}

// Vector-Jacobian product of the full build for one column over layers
// z0 .. z0+63: T_bar[z] = sum_i k_bar[i, z] dk[i, z]/dT[z], M_bar likewise.
// The same expressions as the build, on dual numbers carrying d/dT and d/dM
// (Gibbs energies with their d/dT staged in LDS); each wave accumulates its
// pairs per layer, and the waves' partial sums are added in wave order, so
// the result does not depend on scheduling.
constexpr int VK_VJP_NT = 1024;

__global__ __launch_bounds__(VK_VJP_NT) void k_rates_vjp(VkRatesDev R, int nz,
                                                         const double* __restrict__ Tz,
                                                         const double* __restrict__ Mz,
                                                         int lowt_on,
                                                         const double* __restrict__ kbar,
                                                         double* __restrict__ Tbar,
                                                         double* __restrict__ Mbar) {
#pragma clang fp contract(off)
  extern __shared__ double lds[];
  constexpr int NW = VK_VJP_NT / 64;
  double* lT = lds;                         // (ZT)
  double* lM = lT + VK_RATES_ZT;            // (ZT)
  double* red_t = lM + VK_RATES_ZT;         // (NW, ZT)
  double* red_m = red_t + NW * VK_RATES_ZT; // (NW, ZT)
  double* gv = red_m + NW * VK_RATES_ZT;    // (ni, ZT) g/RT
  double* gt = gv + R.ni * VK_RATES_ZT;     // (ni, ZT) d(g/RT)/dT
  const int tid = threadIdx.x;
  const int z0 = blockIdx.x * VK_RATES_ZT;
  const int zn = min(VK_RATES_ZT, nz - z0);
  if (tid < VK_RATES_ZT) {
    lT[tid] = tid < zn ? Tz[z0 + tid] : 300.0;
    lM[tid] = tid < zn ? Mz[z0 + tid] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < R.ni * VK_RATES_ZT; e += VK_VJP_NT) {
    const int j = e / VK_RATES_ZT, z = e % VK_RATES_ZT;
    const Dual g = gibbs_rt(R.nasa9 + 20 * j, tpow(Dual(lT[z], 1.0, 0.0)));
    gv[e] = g.v;
    gt[e] = g.t;
  }
  __syncthreads();
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  double acc_t = 0.0, acc_m = 0.0;
  if (lane < zn) {
    const int z = z0 + lane;
    const Dual T(lT[lane], 1.0, 0.0), M(lM[lane], 0.0, 1.0);
    const Dual lnT = vlog(T), lc = vlog(VK_GIBBS_CORR * T);
    auto g = [&](int j) {
      return Dual(gv[j * VK_RATES_ZT + lane], gt[j * VK_RATES_ZT + lane], 0.0);
    };
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    using cline_t = const __attribute__((address_space(4))) u32x4*;
    const cline_t lines = (cline_t)R.pairs;
    for (int p = wave; p < R.npairs; p += NW) {
      VkRatePair q;
      u32x4* qv = reinterpret_cast<u32x4*>(&q);
#pragma unroll
      for (int k = 0; k < 8; ++k) qv[k] = lines[8 * p + k];
      if ((q.flags & (VK_PAIR_ZERO_FWD | VK_PAIR_ZERO_REV)) ==
          (VK_PAIR_ZERO_FWD | VK_PAIR_ZERO_REV))
        continue;
      Dual kf, kr;
      build_pair(q, T, lnT, lc, M, lowt_on, R.ni, g, kf, kr);
      const int i = 2 * p + 1;
      const double bf = kbar[(int64_t)i * nz + z];
      acc_t = acc_t + bf * kf.t;
      acc_m = acc_m + bf * kf.m;
      if (i + 1 <= R.nr) {
        const double br = kbar[(int64_t)(i + 1) * nz + z];
        acc_t = acc_t + br * kr.t;
        acc_m = acc_m + br * kr.m;
      }
    }
  }
  red_t[wave * VK_RATES_ZT + lane] = acc_t;
  red_m[wave * VK_RATES_ZT + lane] = acc_m;
  __syncthreads();
  if (wave == 0 && lane < zn) {
    double st = 0.0, sm = 0.0;
    for (int w = 0; w < NW; ++w) {
      st = st + red_t[w * VK_RATES_ZT + lane];
      sm = sm + red_m[w * VK_RATES_ZT + lane];
    }
    if (Tbar) Tbar[z0 + lane] = st;
    if (Mbar) Mbar[z0 + lane] = sm;
  }
}

}  // namespace

int vk_launch_rates_vjp(const VkRatesDev& R, int nz, const double* T, const double* M,
                        int lowt_on, const double* kbar, double* Tbar, double* Mbar,
                        hipStream_t stream) {
  if (nz <= 0) return VK_OK;
  constexpr int NW = VK_VJP_NT / 64;
  const size_t shm = sizeof(double) * VK_RATES_ZT * (2 + 2 * NW + 2 * R.ni);
  static int limit = -1;   // the kernel's dynamic-LDS limit, raised once per process
  if (limit < 0) {
    int dev = 0, per_block = 64 * 1024;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess ||
        hipFuncSetAttribute((const void*)k_rates_vjp, hipFuncAttributeMaxDynamicSharedMemorySize,
                            per_block) != hipSuccess) {
      (void)hipGetLastError();
      per_block = 64 * 1024;
    }
    limit = per_block;
  }
  if (shm > (size_t)limit) {
    vk_set_error("vk_rates_vjp_f64: ni=%d needs %zu bytes of LDS, the limit is %d", R.ni, shm,
                 limit);
    return VK_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(k_rates_vjp, dim3((nz + VK_RATES_ZT - 1) / VK_RATES_ZT), dim3(VK_VJP_NT),
                     shm, stream, R, nz, T, M, lowt_on, kbar, Tbar, Mbar);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_rates_eval(const VkRatesDev& R, int what, int nz, const double* T, const double* M,
                         int lowt_on, const double* in_a, const double* in_b, double* out,
                         hipStream_t stream) {
  if (what == VK_RATES_BUILD) {
    const size_t shm = sizeof(double) * VK_RATES_ZT * (7 + R.ni);
    hipLaunchKernelGGL(k_rates_build_one, dim3((nz + VK_RATES_ZT - 1) / VK_RATES_ZT),
                       dim3(VK_RATES_NT), shm, stream, R, nz, T, M, lowt_on, out);
    VK_CHECK_HIP(hipGetLastError());
    return VK_OK;
  }
  const int rows = (what == VK_RATES_GIBBS) ? R.ni : R.nr + 1;
  const int64_t n = (int64_t)rows * nz;
  if (n <= 0) return VK_OK;
  hipLaunchKernelGGL(k_rates_eval, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, R,
                     what, nz, T, M, lowt_on, in_a, in_b, out);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_rates_build_batch(const VkCol* cols, int ncols, int max_ni, int max_nz,
                                const VkRatesDev* rates, const double* atm, int lowt_on,
                                double* k, hipStream_t stream) {
  if (max_nz <= 0 || ncols <= 0) return VK_OK;
  const size_t shm = sizeof(double) * VK_RATES_ZT * (7 + max_ni);
  hipLaunchKernelGGL(k_rates_build_batch, dim3((max_nz + VK_RATES_ZT - 1) / VK_RATES_ZT, ncols),
                     dim3(VK_RATES_NT), shm, stream, cols, ncols, rates, atm, lowt_on, k);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
