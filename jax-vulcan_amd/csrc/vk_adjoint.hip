// Adjoint of the steady-state residual f(y, theta) = chem_rhs + diffusion
// with respect to its parameters theta, at fixed y and a given lam.
//
// Restates the parameter VJP of _ssi_bwd / _ss_bwd (steady_state_grad.py:
// 214-233, 306-327) in two kernels; both write the NEGATED VJP that the
// implicit-function theorem needs (steady_state_grad.py:226-231, 325-327):
//
//  * k_chem_vjp: one workgroup per (layer, column).  With
//    g_r = sum_prod nu lam - sum_reac nu lam (the stoichiometric projection
//    of lam onto reaction row r) and q_r = prod_slots y^nu (* M if
//    three-body):  cot_k[r, z] = -g_r q_r,
//    cot_M[z] = -sum_{r three-body} g_r k[r, z] q_r / M[z].
//    The M sum uses a fixed reduction tree, so results are reproducible.
//
//  * k_diff_vjp: one thread per element e of the packed atmosphere record.
//    Diffusion is local: element e only enters the rows of the layers next
//    to its own, so the thread evaluates those rows of the diffusion operator
//    (jax_step.py:67-246) in dual numbers seeded at e and sums lam * tangent.
//    Integer-valued leaves (gas mask, mode flags) get a zero cotangent, as
//    the reference's float0 leaves do.  cot_atm has the record layout.
//    The "wide" elements, whose rows are many (a per-layer scalar reaches
//    every species of three layers, ms and alpha every layer of a species),
//    are left to k_diff_vjp_wide: one workgroup per element, the rows spread
//    over its threads and summed by a fixed tree.
#include "vk_common.h"

namespace {

struct Dl {
  double v, d;
};
__device__ inline Dl operator+(Dl a, Dl b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dl operator-(Dl a, Dl b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dl operator*(Dl a, Dl b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dl operator/(Dl a, Dl b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ inline Dl operator-(Dl a) { return {-a.v, -a.d}; }
__device__ inline Dl cst(double c) { return {c, 0.0}; }
// (a > 0) * a and (a < 0) * a: derivative 0 at a == 0, like the reference's masks
__device__ inline Dl posd(Dl a) { return a.v > 0.0 ? a : cst(0.0); }
__device__ inline Dl negd(Dl a) { return a.v < 0.0 ? a : cst(0.0); }

// Record reader that seeds the tangent at one element.
struct Seeded {
  const double* base;
  int64_t seed;
  __device__ Dl operator()(int64_t off) const { return {base[off], off == seed ? 1.0 : 0.0}; }
};

// -1/Hpi[a] + ms g[b] / (Navo kb Ti[a]) + alpha/Ti[a] (Tco[a+1]-Tco[a])/dzi[a]
__device__ inline Dl grav_fac_d(const Seeded& A, const VkAtmOff& o, int i, int a, int b) {
  const Dl nakb = cst(6.02214086e23 * 1.38064852e-16);
  const Dl Ti = A(o.Ti + a);
  return (-cst(1.0) / A(o.Hpi + a) + A(o.ms + i) * A(o.g + b) / (nakb * Ti)) +
         A(o.alpha + i) / Ti * (A(o.Tco + a + 1) - A(o.Tco + a)) / A(o.dzi + a);
}

// Row (z, i) of the diffusion operator with the flux boundary terms,
// jax_step.py:150-246, as a dual number.
__device__ Dl diff_row_d(const Seeded& A, const VkAtmOff& o, int nz, int ni, int z, int i,
                         const double* yc, const double* ysum, double use_vm, double use_set,
                         bool use_top, bool use_bot) {
  const Dl one = cst(1.0), two = cst(2.0), half = cst(0.5);
  const double ysm = z > 0 ? ysum[z - 1] : 0.0, ys0 = ysum[z], ysp = z < nz - 1 ? ysum[z + 1] : 0.0;
  const Dl Ysm = cst(ysm), Ys0 = cst(ys0), Ysp = cst(ysp);
  const Dl uvm = cst(use_vm), nvm = cst(1.0 - use_vm), ust = cst(use_set);
  Dl At, Bt = cst(0.0), Ct = cst(0.0);
  if (z == 0) {
    const Dl d0 = A(o.dzi), K0 = A(o.Kzz), v0 = A(o.vz), D0 = A(o.Dzz + i);
    const Dl eA = -one / d0 * (K0 / d0) * (Ysp + Ys0) / two / Ys0 - posd(v0) / d0;
    const Dl eB = one / d0 * (K0 / d0) * (Ysp + Ys0) / two / Ysp - negd(v0) / d0;
    const Dl g0 = one / d0 * D0 / two * grav_fac_d(A, o, i, 0, 0);
    const Dl vm0 = A(o.vm + i), vs0 = A(o.vs + i);
    const Dl bA = (nvm * g0 + uvm * (-posd(vm0) / d0)) + ust * (-posd(vs0) / d0);
    const Dl bB = (nvm * g0 + uvm * (-negd(vm0) / d0)) + ust * (-negd(vs0) / d0);
    const Dl mA = -one / d0 * (D0 / d0) * (Ysp + Ys0) / two / Ys0 + bA;
    const Dl mB = one / d0 * (D0 / d0) * (Ysp + Ys0) / two / Ysp + bB;
    At = eA + mA;
    Bt = eB + mB;
  } else if (z == nz - 1) {
    const int t = nz - 2;
    const Dl dt = A(o.dzi + t), Kt = A(o.Kzz + t), vt = A(o.vz + t), Dt = A(o.Dzz + (int64_t)t * ni + i);
    const Dl eA = -one / dt * (Kt / dt) * (Ys0 + Ysm) / two / Ys0 + negd(vt) / dt;
    const Dl eC = one / dt * (Kt / dt) * (Ys0 + Ysm) / two / Ysm + posd(vt) / dt;
    const Dl gt = -one / dt * Dt / two * grav_fac_d(A, o, i, t, nz - 1);
    const Dl vmt = A(o.vm + (int64_t)(nz - 1) * ni + i), vst = A(o.vs + (int64_t)t * ni + i);
    const Dl bA = (nvm * gt + uvm * (negd(vmt) / dt)) + ust * (negd(vst) / dt);
    const Dl bC = (nvm * gt + uvm * (posd(vmt) / dt)) + ust * (posd(vst) / dt);
    const Dl mA = -one / dt * (Dt / dt) * (Ys0 + Ysm) / two / Ys0 + bA;
    const Dl mC = one / dt * (Dt / dt) * (Ys0 + Ysm) / two / Ysm + bC;
    At = eA + mA;
    Ct = eC + mC;
  } else {
    const Dl dm = A(o.dzi + z - 1), dj = A(o.dzi + z);
    const Dl dza = half * (dm + dj);
    const Dl Kj = A(o.Kzz + z), Km = A(o.Kzz + z - 1), vj = A(o.vz + z), vmz = A(o.vz + z - 1);
    const Dl eA = -one / dza * (Kj / dj * (Ysp + Ys0) / two + Km / dm * (Ys0 + Ysm) / two) / Ys0 -
                  (posd(vj) - negd(vmz)) / dza;
    const Dl eB = one / dza * Kj / dj * (Ysp + Ys0) / two / Ysp - negd(vj) / dza;
    const Dl eC = one / dza * Km / dm * (Ys0 + Ysm) / two / Ysm + posd(vmz) / dza;
    const Dl Dj = A(o.Dzz + (int64_t)z * ni + i), Djm = A(o.Dzz + (int64_t)(z - 1) * ni + i);
    const Dl inv2 = one / (two * dza);
    const Dl gA = inv2 * (Dj * grav_fac_d(A, o, i, z, z) - Djm * grav_fac_d(A, o, i, z - 1, z));
    const Dl gB = inv2 * Dj * grav_fac_d(A, o, i, z, z + 1);
    const Dl gC = -inv2 * Djm * grav_fac_d(A, o, i, z - 1, z - 1);
    const Dl vmj = A(o.vm + (int64_t)z * ni + i), vmm = A(o.vm + (int64_t)(z - 1) * ni + i);
    const Dl vsj = A(o.vs + (int64_t)z * ni + i), vsm = A(o.vs + (int64_t)(z - 1) * ni + i);
    const Dl Avm = -(posd(vmj) - negd(vmm)) / dza, Bvm = -negd(vmj) / dza, Cvm = posd(vmm) / dza;
    const Dl Avs = -(posd(vsj) - negd(vsm)) / dza, Bvs = -negd(vsj) / dza, Cvs = posd(vsm) / dza;
    const Dl Ax = (nvm * gA + uvm * Avm) + ust * Avs;
    const Dl Bx = (nvm * gB + uvm * Bvm) + ust * Bvs;
    const Dl Cx = (nvm * gC + uvm * Cvm) + ust * Cvs;
    const Dl mA = -one / dza * (Dj / dj * (Ysp + Ys0) / two + Djm / dm * (Ys0 + Ysm) / two) / Ys0 + Ax;
    const Dl mB = one / dza * Dj / dj * (Ysp + Ys0) / two / Ysp + Bx;
    const Dl mC = one / dza * Djm / dm * (Ys0 + Ysm) / two / Ysm + Cx;
    At = eA + mA;
    Bt = eB + mB;
    Ct = eC + mC;
  }
  const Dl y0 = cst(yc[(int64_t)z * ni + i]);
  Dl d = At * y0;
  if (z < nz - 1) d = d + Bt * cst(yc[(int64_t)(z + 1) * ni + i]);
  if (z > 0) d = d + Ct * cst(yc[(int64_t)(z - 1) * ni + i]);
  if (z == nz - 1 && use_top) d = d + A(o.top_flux + i) / A(o.dzi + nz - 2);
  if (z == 0 && use_bot) d = d + (A(o.bot_flux + i) - y0 * A(o.bot_vdep + i)) / A(o.dzi);
  return d;
}

}  // namespace

// The layers and species element e of the record reaches (z in [z_lo, z_hi),
// i in [i_lo, i_hi)); true for the wide elements (see above).
__device__ inline bool diff_window(const VkAtmOff& o, int nz, int ni, int64_t e, int& z_lo,
                                   int& z_hi, int& i_lo, int& i_hi) {
  bool wide = false;
  if (e < o.Tco + nz || (e >= o.g && e < o.g + nz)) {  // per-layer scalars
    int m;
    if (e < o.Tco) m = (int)((e - o.Kzz) % (nz - 1));
    else if (e < o.g) m = (int)(e - o.Tco);
    else m = (int)(e - o.g);
    z_lo = m - 1; z_hi = m + 2; i_lo = 0; i_hi = ni;
    wide = true;
  } else if (e < o.top_flux) {  // ms, alpha: every layer of species i
    const int i = (int)((e - o.ms) % ni);
    z_lo = 0; z_hi = nz; i_lo = i; i_hi = i + 1;
    wide = true;
  } else if (e < o.bot_flux) {  // top_flux
    const int i = (int)(e - o.top_flux);
    z_lo = nz - 1; z_hi = nz; i_lo = i; i_hi = i + 1;
  } else if (e < o.gas) {  // bot_flux, bot_vdep
    const int i = (int)((e - o.bot_flux) % ni);
    z_lo = 0; z_hi = 1; i_lo = i; i_hi = i + 1;
  } else {  // Dzz (nz-1, ni), vm (nz, ni), vs (nz-1, ni)
    int64_t r;
    if (e < o.vm) r = e - o.Dzz;
    else if (e < o.vs) r = e - o.vm;
    else r = e - o.vs;
    const int m = (int)(r / ni), i = (int)(r % ni);
    z_lo = m - 1; z_hi = m + 2; i_lo = i; i_hi = i + 1;
  }
  z_lo = z_lo < 0 ? 0 : z_lo;
  z_hi = z_hi > nz ? nz : z_hi;
  return wide;
}

// ysum over the gas species (jax_step.py:165-167) of layers [z0, z1), one
// layer per thread, summed in species order as the chemistry kernels do.
__device__ inline void ysum_layers(const double* rec, const VkAtmOff& o, const double* yc, int ni,
                                   int z0, int z1, double* ysum) {
  for (int z = z0 + (int)threadIdx.x; z < z1; z += VK_NT) {
    double s = 0.0;
    for (int i = 0; i < ni; ++i)
      if (rec[o.gas + i] != 0.0) s = s + yc[(int64_t)z * ni + i];
    ysum[z] = s;
  }
}

__global__ __launch_bounds__(VK_NT) void k_diff_vjp(const VkCol* __restrict__ cols, int ncols,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ atm,
                                                    const double* __restrict__ lam,
                                                    double* __restrict__ cot_atm) {
  extern __shared__ double ysum[];
  const int c = blockIdx.y;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  const int nz = col.nz, ni = col.ni;
  const VkAtmOff o = vk_atm_offsets(nz, ni);
  const double* rec = atm + col.off_a;
  const double* yc = y + col.off_s;
  const double* lc = lam + col.off_s;
  // the ysum of the layers this workgroup's elements reach
  {
    const int64_t e0 = (int64_t)blockIdx.x * VK_NT;
    int z0 = 0, z1 = nz;
    if (e0 >= o.Dzz) {   // Dzz, vm, vs rows: layers around the block's rows
      const int64_t r0 = e0 < o.vm ? e0 - o.Dzz : (e0 < o.vs ? e0 - o.vm : e0 - o.vs);
      const int64_t r1 = r0 + VK_NT;
      z0 = (int)(r0 / ni) - 2;
      z1 = (int)(r1 / ni) + 3;
      if ((e0 < o.vm && e0 + VK_NT > o.vm) || (e0 < o.vs && e0 + VK_NT > o.vs)) {
        z0 = 0;   // the block straddles two fields
        z1 = nz;
      }
      z0 = z0 < 0 ? 0 : z0;
      z1 = z1 > nz ? nz : z1;
    }
    ysum_layers(rec, o, yc, ni, z0, z1, ysum);
  }
  __syncthreads();
  const int64_t e = (int64_t)blockIdx.x * VK_NT + threadIdx.x;
  if (e >= o.total) return;
  double* out = cot_atm + col.off_a;
  if (e >= o.M && e < o.M + nz) return;  // M enters the chemistry only: k_chem_vjp writes it
  if ((e >= o.gas && e < o.gas + ni) || e >= o.flags) {
    out[e] = 0.0;
    return;
  }
  const double use_vm = rec[o.flags + 0] != 0.0 ? 1.0 : 0.0;
  const double use_set = rec[o.flags + 1] != 0.0 ? 1.0 : 0.0;
  const bool use_top = rec[o.flags + 2] != 0.0, use_bot = rec[o.flags + 3] != 0.0;
  const Seeded A{rec, e};
  // the layer window and species range element e can reach
  int z_lo, z_hi, i_lo, i_hi;
  if (diff_window(o, nz, ni, e, z_lo, z_hi, i_lo, i_hi)) return;  // k_diff_vjp_wide
  double acc = 0.0;
  for (int z = z_lo; z < z_hi; ++z)
    for (int i = i_lo; i < i_hi; ++i) {
      const Dl d = diff_row_d(A, o, nz, ni, z, i, yc, ysum, use_vm, use_set, use_top, use_bot);
      acc = acc + lc[(int64_t)z * ni + i] * d.d;
    }
  out[e] = -acc;
}

// One workgroup per wide element: the Kzz .. g block and the ms, alpha
// block of the record, in that order along blockIdx.x.
__global__ __launch_bounds__(VK_NT) void k_diff_vjp_wide(const VkCol* __restrict__ cols,
                                                         int ncols, const double* __restrict__ y,
                                                         const double* __restrict__ atm,
                                                         const double* __restrict__ lam,
                                                         double* __restrict__ cot_atm) {
  extern __shared__ double ysum[];
  __shared__ double red[VK_NT / 64];
  const int c = blockIdx.y;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  const int nz = col.nz, ni = col.ni;
  const VkAtmOff o = vk_atm_offsets(nz, ni);
  const int64_t n1 = o.M - o.Kzz, n2 = o.top_flux - o.ms;
  const int64_t w = blockIdx.x;
  if (w >= n1 + n2) return;
  const int64_t e = w < n1 ? o.Kzz + w : o.ms + (w - n1);
  const double* rec = atm + col.off_a;
  const double* yc = y + col.off_s;
  const double* lc = lam + col.off_s;
  int z_lo, z_hi, i_lo, i_hi;
  diff_window(o, nz, ni, e, z_lo, z_hi, i_lo, i_hi);
  ysum_layers(rec, o, yc, ni, z_lo > 0 ? z_lo - 1 : 0, z_hi < nz ? z_hi + 1 : nz, ysum);
  __syncthreads();
  const double use_vm = rec[o.flags + 0] != 0.0 ? 1.0 : 0.0;
  const double use_set = rec[o.flags + 1] != 0.0 ? 1.0 : 0.0;
  const bool use_top = rec[o.flags + 2] != 0.0, use_bot = rec[o.flags + 3] != 0.0;
  const Seeded A{rec, e};
  const int nsp = i_hi - i_lo, nrow = (z_hi - z_lo) * nsp;
  double acc = 0.0;
  for (int q = threadIdx.x; q < nrow; q += VK_NT) {
    const int z = z_lo + q / nsp, i = i_lo + q % nsp;
    const Dl d = diff_row_d(A, o, nz, ni, z, i, yc, ysum, use_vm, use_set, use_top, use_bot);
    acc = acc + lc[(int64_t)z * ni + i] * d.d;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) cot_atm[col.off_a + e] = -((red[0] + red[1]) + (red[2] + red[3]));
}

__global__ __launch_bounds__(VK_NT) void k_chem_vjp(const VkCol* __restrict__ cols, int ncols,
                                                    const VkNetDev* __restrict__ nets,
                                                    const double* __restrict__ y,
                                                    const double* __restrict__ atm,
                                                    const double* __restrict__ k,
                                                    const double* __restrict__ lam,
                                                    double* __restrict__ cot_k,
                                                    double* __restrict__ cot_atm) {
  extern __shared__ double smem[];
  __shared__ double red[VK_NT / 64];
  const int c = blockIdx.y, z = blockIdx.x;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  const int nz = col.nz, ni = col.ni, tid = threadIdx.x;
  if (z >= nz) return;
  const VkNetDev net = nets[col.net];
  const int S = net.S, T = net.T;
  double* yz = smem;           // ni + 1 (PAD = 1)
  double* lz = smem + ni + 1;  // ni + 1 (PAD = 0)
  for (int i = tid; i <= ni; i += VK_NT) {
    yz[i] = i < ni ? y[col.off_s + (int64_t)z * ni + i] : 1.0;
    lz[i] = i < ni ? lam[col.off_s + (int64_t)z * ni + i] : 0.0;
  }
  __syncthreads();
  const VkAtmOff ao = vk_atm_offsets(nz, ni);
  const double Mz = atm[col.off_a + ao.M + z];
  double mpart = 0.0;
  for (int r = tid; r < S; r += VK_NT) {
    double q = 1.0, g = 0.0;
    for (int s = 0; s < T; ++s) {
      const int sp = net.r_idx[r * T + s];
      const double st = net.r_st[r * T + s];
      if (st > 0.0) q = q * vk_ipow(yz[sp], st);
      g = g - st * lz[sp];
    }
    for (int s = 0; s < T; ++s) g = g + net.p_st[r * T + s] * lz[net.p_idx[r * T + s]];
    const bool tb = net.is3[r] != 0;
    const double qm = tb ? q * Mz : q;
    const int64_t e = col.off_k + (int64_t)r * nz + z;
    if (cot_k) cot_k[e] = -(g * qm);
    if (tb) mpart += g * k[e] * q;
  }
  if (!cot_atm) return;
  for (int off = 32; off > 0; off >>= 1) mpart += __shfl_xor(mpart, off, 64);
  if ((tid & 63) == 0) red[tid >> 6] = mpart;
  __syncthreads();
  if (tid == 0) cot_atm[col.off_a + ao.M + z] = -((red[0] + red[1]) + (red[2] + red[3]));
}

int vk_launch_ss_vjp(const VkCol* cols, int ncols, int max_nz, int max_ni, int64_t max_rec,
                     const VkNetDev* nets, const double* y, const double* atm, const double* k,
                     const double* lam, double* cot_k, double* cot_atm, hipStream_t st) {
  if (cot_atm) {
    const int gx = (int)((max_rec + VK_NT - 1) / VK_NT);
    hipLaunchKernelGGL(k_diff_vjp, dim3(gx, ncols), dim3(VK_NT), sizeof(double) * (size_t)max_nz,
                       st, cols, ncols, y, atm, lam, cot_atm);
    VK_CHECK_HIP(hipGetLastError());
    // wide elements: Kzz, dzi, vz, Hpi, Ti (nz-1 each), Tco, g (nz), ms, alpha (ni)
    const int gw = 5 * (max_nz - 1) + 2 * max_nz + 2 * max_ni;
    hipLaunchKernelGGL(k_diff_vjp_wide, dim3(gw, ncols), dim3(VK_NT),
                       sizeof(double) * (size_t)max_nz, st, cols, ncols, y, atm, lam, cot_atm);
    VK_CHECK_HIP(hipGetLastError());
  }
  const size_t shmem = sizeof(double) * 2 * (size_t)(max_ni + 1);
  hipLaunchKernelGGL(k_chem_vjp, dim3(max_nz, ncols), dim3(VK_NT), shmem, st, cols, ncols, nets,
                     y, atm, k, lam, cot_k, cot_atm);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
