// Chemistry right-hand side, analytical Jacobian and eddy/molecular
// diffusion for one (layer, column) per workgroup.
//
// Restates chem.chem_rhs_per_layer (chem.py:83-125),
// chem.chem_jac_analytical_per_layer (chem.py:144-208) and the diffusion
// operator of jax_step.py:67-246.  Sums run over precomputed CSR lists in a
// fixed order, with FP contraction off, so they repeat bit for bit from run
// to run and from batch to batch (no atomics).
#include "vk_common.h"

namespace {

// Kernel modes.
enum {
  M_CHEM_RHS = 1,  // out0 = chem rhs (natural nz x ni)
  M_CHEM_JAC = 2,  // out0 = chem Jacobian (natural nz x ni x ni)
  M_LHS = 3,       // Ros2 system at y: D tiles, rhs_y, -sup, -sub
  M_RHS2 = 4,      // rhs2 = f(yk2) - 2/(gamma dt) k1
  M_FTOT = 5,      // out0 = chem + diffusion
  M_JY = 6,        // J_y = chem_J + diag(A+Ai) blocks (natural), sup, sub
};

struct LayerArgs {
  const VkCol* cols;
  int ncols;
  const VkNetDev* nets;
  const double* y;       // state the operator is evaluated at (off_s)
  const double* M;       // per-layer M at off_s/ni*... or null: atm record
  const double* atm;     // packed atmosphere records (off_a)
  const double* k;       // (S, nz) rate tables (off_k)
  const double* dt;      // per column
  const uint8_t* fix;    // optional fix mask (off_s)
  const double* k1;      // M_RHS2
  double* out0;          // rhs-like output (off_s) or natural blocks
  double* out_blk;       // M_LHS: tile blocks at off_b
  double* out_sup;       // (nz-1, ni) at off_s
  double* out_sub;
  const int* active;
  int mode;
  int k_layer_major;     // k holds (nz, S) per column (a transposed copy)
  const double* mstat;   // optional molecular-diffusion static factors (k_mol_static)
  int64_t mstride;       // component stride of mstat
};

constexpr double GAMMA = 1.7071067811865475;  // 1 + 1/sqrt(2)

__device__ inline double posp(double v) { return v > 0.0 ? v : 0.0; }
__device__ inline double negp(double v) { return v < 0.0 ? v : 0.0; }

// Diffusion coefficients of layer z: eddy (A, B, C) and, for species i, the
// molecular (Ai, Bi, Ci), following jax_step.py:67-220 term by term.
struct DiffCoef {
  double A, B, C;
};

__device__ inline DiffCoef eddy_coef(int z, int nz, const double* Kzz, const double* dzi,
                                     const double* vz, double ysm, double ys0, double ysp) {
#pragma clang fp contract(off)
  DiffCoef d;
  if (z == 0) {
    d.A = -1.0 / dzi[0] * (Kzz[0] / dzi[0]) * (ysp + ys0) / 2.0 / ys0 - posp(vz[0]) / dzi[0];
    d.B = 1.0 / dzi[0] * (Kzz[0] / dzi[0]) * (ysp + ys0) / 2.0 / ysp - negp(vz[0]) / dzi[0];
    d.C = 0.0;
  } else if (z == nz - 1) {
    const int t = nz - 2;
    d.A = -1.0 / dzi[t] * (Kzz[t] / dzi[t]) * (ys0 + ysm) / 2.0 / ys0 + negp(vz[t]) / dzi[t];
    d.C = 1.0 / dzi[t] * (Kzz[t] / dzi[t]) * (ys0 + ysm) / 2.0 / ysm + posp(vz[t]) / dzi[t];
    d.B = 0.0;
  } else {
    const double dza = 0.5 * (dzi[z - 1] + dzi[z]);
    d.A = -1.0 / dza *
              (Kzz[z] / dzi[z] * (ysp + ys0) / 2.0 + Kzz[z - 1] / dzi[z - 1] * (ys0 + ysm) / 2.0) /
              ys0 -
          (posp(vz[z]) - negp(vz[z - 1])) / dza;
    d.B = 1.0 / dza * Kzz[z] / dzi[z] * (ysp + ys0) / 2.0 / ysp - negp(vz[z]) / dza;
    d.C = 1.0 / dza * Kzz[z - 1] / dzi[z - 1] * (ys0 + ysm) / 2.0 / ysm + posp(vz[z - 1]) / dza;
  }
  return d;
}

struct AtmView {
  const double *Kzz, *dzi, *vz, *Hpi, *Ti, *Tco, *g, *M;
  const double *ms, *alpha, *top_flux, *bot_flux, *bot_vdep, *gas;
  const double *Dzz, *vm, *vs, *flags;
};

__device__ inline AtmView atm_view(const double* base, int nz, int ni) {
  const VkAtmOff o = vk_atm_offsets(nz, ni);
  AtmView v;
  v.Kzz = base + o.Kzz; v.dzi = base + o.dzi; v.vz = base + o.vz; v.Hpi = base + o.Hpi;
  v.Ti = base + o.Ti; v.Tco = base + o.Tco; v.g = base + o.g; v.M = base + o.M;
  v.ms = base + o.ms; v.alpha = base + o.alpha; v.top_flux = base + o.top_flux;
  v.bot_flux = base + o.bot_flux; v.bot_vdep = base + o.bot_vdep; v.gas = base + o.gas;
  v.Dzz = base + o.Dzz; v.vm = base + o.vm; v.vs = base + o.vs; v.flags = base + o.flags;
  return v;
}

// gravity/thermal-diffusion factor of compute_diff_grav (jax_step.py:80-87):
// -1/Hpi[a] + ms g[b] / (Navo kb Ti[a]) + alpha/Ti[a] (Tco[a+1]-Tco[a])/dzi[a]
__device__ inline double grav_fac(const AtmView& v, int i, int a, int b) {
#pragma clang fp contract(off)
  const double NAKB = 6.02214086e23 * 1.38064852e-16;
  return -1.0 / v.Hpi[a] + v.ms[i] * v.g[b] / (NAKB * v.Ti[a]) +
         v.alpha[i] / v.Ti[a] * (v.Tco[a + 1] - v.Tco[a]) / v.dzi[a];
}

// Molecular coefficients of species i at layer z (jax_step.py:80-147, 195-218).
__device__ inline DiffCoef mol_coef(int z, int nz, int ni, int i, const AtmView& v, double ysm,
                                    double ys0, double ysp) {
#pragma clang fp contract(off)
  const double use_vm = v.flags[0] != 0.0 ? 1.0 : 0.0;
  const double use_set = v.flags[1] != 0.0 ? 1.0 : 0.0;
  const double* dzi = v.dzi;
  DiffCoef d;
  if (z == 0) {
    const double D0 = v.Dzz[i];
    const double g0 = 1.0 / dzi[0] * D0 / 2.0 * grav_fac(v, i, 0, 0);
    const double vm0 = v.vm[i], vs0 = v.vs[i];
    const double bA = (1.0 - use_vm) * g0 + use_vm * (-posp(vm0) / dzi[0]) + use_set * (-posp(vs0) / dzi[0]);
    const double bB = (1.0 - use_vm) * g0 + use_vm * (-negp(vm0) / dzi[0]) + use_set * (-negp(vs0) / dzi[0]);
    d.A = -1.0 / dzi[0] * (D0 / dzi[0]) * (ysp + ys0) / 2.0 / ys0 + bA;
    d.B = 1.0 / dzi[0] * (D0 / dzi[0]) * (ysp + ys0) / 2.0 / ysp + bB;
    d.C = 0.0;
  } else if (z == nz - 1) {
    const int t = nz - 2;
    const double Dt = v.Dzz[(int64_t)t * ni + i];
    const double gt = -1.0 / dzi[t] * Dt / 2.0 *
                      (-1.0 / v.Hpi[t] + v.ms[i] * v.g[nz - 1] / (6.02214086e23 * 1.38064852e-16 * v.Ti[t]) +
                       v.alpha[i] / v.Ti[t] * (v.Tco[nz - 1] - v.Tco[nz - 2]) / dzi[t]);
    const double vmt = v.vm[(int64_t)(nz - 1) * ni + i], vst = v.vs[(int64_t)t * ni + i];
    const double bA = (1.0 - use_vm) * gt + use_vm * (negp(vmt) / dzi[t]) + use_set * (negp(vst) / dzi[t]);
    const double bC = (1.0 - use_vm) * gt + use_vm * (posp(vmt) / dzi[t]) + use_set * (posp(vst) / dzi[t]);
    d.A = -1.0 / dzi[t] * (Dt / dzi[t]) * (ys0 + ysm) / 2.0 / ys0 + bA;
    d.C = 1.0 / dzi[t] * (Dt / dzi[t]) * (ys0 + ysm) / 2.0 / ysm + bC;
    d.B = 0.0;
  } else {
    const double dza = 0.5 * (dzi[z - 1] + dzi[z]);
    const double Dj = v.Dzz[(int64_t)z * ni + i], Djm = v.Dzz[(int64_t)(z - 1) * ni + i];
    const double inv2 = 1.0 / (2.0 * dza);
    const double gA = inv2 * (Dj * grav_fac(v, i, z, z) - Djm * grav_fac(v, i, z - 1, z));
    const double gB = inv2 * Dj * grav_fac(v, i, z, z + 1);
    const double gC = -inv2 * Djm * grav_fac(v, i, z - 1, z - 1);
    const double vmj = v.vm[(int64_t)z * ni + i], vmm = v.vm[(int64_t)(z - 1) * ni + i];
    const double vsj = v.vs[(int64_t)z * ni + i], vsm = v.vs[(int64_t)(z - 1) * ni + i];
    const double Avm = -(posp(vmj) - negp(vmm)) / dza, Bvm = -negp(vmj) / dza, Cvm = posp(vmm) / dza;
    const double Avs = -(posp(vsj) - negp(vsm)) / dza, Bvs = -negp(vsj) / dza, Cvs = posp(vsm) / dza;
    const double Ax = ((1.0 - use_vm) * gA + use_vm * Avm) + use_set * Avs;
    const double Bx = ((1.0 - use_vm) * gB + use_vm * Bvm) + use_set * Bvs;
    const double Cx = ((1.0 - use_vm) * gC + use_vm * Cvm) + use_set * Cvs;
    d.A = -1.0 / dza *
              (Dj / dzi[z] * (ysp + ys0) / 2.0 + Djm / dzi[z - 1] * (ys0 + ysm) / 2.0) / ys0 +
          Ax;
    d.B = 1.0 / dza * Dj / dzi[z] * (ysp + ys0) / 2.0 / ysp + Bx;
    d.C = 1.0 / dza * Djm / dzi[z - 1] * (ys0 + ysm) / 2.0 / ysm + Cx;
  }
  return d;
}

// The y-independent factors of mol_coef for species i at layer z, each the
// value mol_coef forms on its way (same operations, same order), so that
// mol_coef_pre below reproduces mol_coef bit for bit:
//   m[0]  A prefix: -1/dzi0 * (D0/dzi0) (z = 0), -1/dzi_t * (Dt/dzi_t)
//         (z = nz-1), -1/dza (inside)
//   m[1], m[2]  inside: Dj/dzi[z], Djm/dzi[z-1]
//   m[3]  B prefix: 1/dzi0 * (D0/dzi0) (z = 0), (1/dza * Dj)/dzi[z] (inside)
//   m[4]  C prefix: 1/dzi_t * (Dt/dzi_t) (z = nz-1), (1/dza * Djm)/dzi[z-1]
//   m[5], m[6], m[7]  the additive A, B, C terms (gravity, vm, settling)
__device__ inline void mol_static(int z, int nz, int ni, int i, const AtmView& v, double (&m)[8]) {
#pragma clang fp contract(off)
  const double use_vm = v.flags[0] != 0.0 ? 1.0 : 0.0;
  const double use_set = v.flags[1] != 0.0 ? 1.0 : 0.0;
  const double* dzi = v.dzi;
#pragma unroll
  for (int c = 0; c < 8; ++c) m[c] = 0.0;
  if (z == 0) {
    const double D0 = v.Dzz[i];
    const double g0 = 1.0 / dzi[0] * D0 / 2.0 * grav_fac(v, i, 0, 0);
    const double vm0 = v.vm[i], vs0 = v.vs[i];
    m[5] = (1.0 - use_vm) * g0 + use_vm * (-posp(vm0) / dzi[0]) + use_set * (-posp(vs0) / dzi[0]);
    m[6] = (1.0 - use_vm) * g0 + use_vm * (-negp(vm0) / dzi[0]) + use_set * (-negp(vs0) / dzi[0]);
    m[0] = -1.0 / dzi[0] * (D0 / dzi[0]);
    m[3] = 1.0 / dzi[0] * (D0 / dzi[0]);
  } else if (z == nz - 1) {
    const int t = nz - 2;
    const double Dt = v.Dzz[(int64_t)t * ni + i];
    const double gt = -1.0 / dzi[t] * Dt / 2.0 *
                      (-1.0 / v.Hpi[t] + v.ms[i] * v.g[nz - 1] / (6.02214086e23 * 1.38064852e-16 * v.Ti[t]) +
                       v.alpha[i] / v.Ti[t] * (v.Tco[nz - 1] - v.Tco[nz - 2]) / dzi[t]);
    const double vmt = v.vm[(int64_t)(nz - 1) * ni + i], vst = v.vs[(int64_t)t * ni + i];
    m[5] = (1.0 - use_vm) * gt + use_vm * (negp(vmt) / dzi[t]) + use_set * (negp(vst) / dzi[t]);
    m[7] = (1.0 - use_vm) * gt + use_vm * (posp(vmt) / dzi[t]) + use_set * (posp(vst) / dzi[t]);
    m[0] = -1.0 / dzi[t] * (Dt / dzi[t]);
    m[4] = 1.0 / dzi[t] * (Dt / dzi[t]);
  } else {
    const double dza = 0.5 * (dzi[z - 1] + dzi[z]);
    const double Dj = v.Dzz[(int64_t)z * ni + i], Djm = v.Dzz[(int64_t)(z - 1) * ni + i];
    const double inv2 = 1.0 / (2.0 * dza);
    const double gA = inv2 * (Dj * grav_fac(v, i, z, z) - Djm * grav_fac(v, i, z - 1, z));
    const double gB = inv2 * Dj * grav_fac(v, i, z, z + 1);
    const double gC = -inv2 * Djm * grav_fac(v, i, z - 1, z - 1);
    const double vmj = v.vm[(int64_t)z * ni + i], vmm = v.vm[(int64_t)(z - 1) * ni + i];
    const double vsj = v.vs[(int64_t)z * ni + i], vsm = v.vs[(int64_t)(z - 1) * ni + i];
    const double Avm = -(posp(vmj) - negp(vmm)) / dza, Bvm = -negp(vmj) / dza, Cvm = posp(vmm) / dza;
    const double Avs = -(posp(vsj) - negp(vsm)) / dza, Bvs = -negp(vsj) / dza, Cvs = posp(vsm) / dza;
    m[5] = ((1.0 - use_vm) * gA + use_vm * Avm) + use_set * Avs;
    m[6] = ((1.0 - use_vm) * gB + use_vm * Bvm) + use_set * Bvs;
    m[7] = ((1.0 - use_vm) * gC + use_vm * Cvm) + use_set * Cvs;
    m[0] = -1.0 / dza;
    m[1] = Dj / dzi[z];
    m[2] = Djm / dzi[z - 1];
    m[3] = 1.0 / dza * Dj / dzi[z];
    m[4] = 1.0 / dza * Djm / dzi[z - 1];
  }
}

// mol_coef from the factors of mol_static (component c at m[c * st]).
__device__ inline DiffCoef mol_coef_pre(int z, int nz, const double* m, int64_t st, double ysm,
                                        double ys0, double ysp) {
#pragma clang fp contract(off)
  DiffCoef d;
  if (z == 0) {
    d.A = m[0] * (ysp + ys0) / 2.0 / ys0 + m[5 * st];
    d.B = m[3 * st] * (ysp + ys0) / 2.0 / ysp + m[6 * st];
    d.C = 0.0;
  } else if (z == nz - 1) {
    d.A = m[0] * (ys0 + ysm) / 2.0 / ys0 + m[5 * st];
    d.C = m[4 * st] * (ys0 + ysm) / 2.0 / ysm + m[7 * st];
    d.B = 0.0;
  } else {
    d.A = m[0] * (m[st] * (ysp + ys0) / 2.0 + m[2 * st] * (ys0 + ysm) / 2.0) / ys0 + m[5 * st];
    d.B = m[3 * st] * (ysp + ys0) / 2.0 / ysp + m[6 * st];
    d.C = m[4 * st] * (ys0 + ysm) / 2.0 / ysm + m[7 * st];
  }
  return d;
}

// Apply f to the records of one lane stream (records rec[k * VK_CHEM_NT],
// k < count, count a multiple of VK_REC_GROUP), in order.  Records are
// loaded a group at a time, and the next group is requested before the
// current one is worked through, so the loads overlap the work instead of
// waiting behind its branches.
//
// Every record has a term (a term-less one points at a zero slot), so the
// terms of a whole group, term(rec) = c * table[item], are formed first, all
// their table reads in flight together, and close(rec, term) then adds each
// to the running sum in record order and stores closed sums.
//
// EVEN: the last group reads itself again, so every pass issues the same
// eight loads and the wait for the current group leaves the next group's
// loads in flight (with a pass that issued none, that wait drains every
// load).  It pays in the walks whose closes write LDS (the loss/production
// sums); the Jacobian walk, whose closes store to memory, measured faster
// without it.
template <bool EVEN, class Term, class Close>
__device__ __forceinline__ void walk_records(const VkRec* rec, int count, Term&& term,
                                             Close&& close) {
  VkRec cur[VK_REC_GROUP], nxt[VK_REC_GROUP];
#pragma unroll
  for (int g = 0; g < VK_REC_GROUP; ++g) cur[g] = rec[(int64_t)g * VK_CHEM_NT];
  for (int k0 = 0; k0 < count; k0 += VK_REC_GROUP) {
    if constexpr (EVEN) {
      const int kn = k0 + VK_REC_GROUP < count ? k0 + VK_REC_GROUP : k0;
#pragma unroll
      for (int g = 0; g < VK_REC_GROUP; ++g) nxt[g] = rec[(int64_t)(kn + g) * VK_CHEM_NT];
    } else if (k0 + VK_REC_GROUP < count) {
#pragma unroll
      for (int g = 0; g < VK_REC_GROUP; ++g)
        nxt[g] = rec[(int64_t)(k0 + VK_REC_GROUP + g) * VK_CHEM_NT];
    }
    double t[VK_REC_GROUP];
#pragma unroll
    for (int g = 0; g < VK_REC_GROUP; ++g) t[g] = term(cur[g]);
#pragma unroll
    for (int g = 0; g < VK_REC_GROUP; ++g) close(cur[g], t[g]);
#pragma unroll
    for (int g = 0; g < VK_REC_GROUP; ++g) cur[g] = nxt[g];
  }
}

}  // namespace

// Dynamic shared memory: rate[S] + drate[S*T] (Jacobian modes) + 3 y rows
// + the loss / production chunk sums.
// NT threads per workgroup: VK_CHEM_NT for the Jacobian modes, whose LDS
// (rates and their derivatives) allows two workgroups per CU, and half that
// for the others; each thread walks the record streams of lanes tid,
// tid + NT, ... of the VK_CHEM_NT-lane tables.
// The 256-thread (right-hand-side) launches are compiled for six waves per
// SIMD (80 VGPRs; what spills are scalar registers, into vector lanes), so
// five workgroups fit a CU as their LDS allows, not four: the stage-2 RHS
// went from 3.79 to 3.33 ms at 1024 columns.  The 512-thread Jacobian
// launches keep their registers: their LDS allows two per CU either way,
// and a variant that formed the derivatives after the chunk sums, in their
// space, to fit three ran slower (10.6 against 9.9 ms).
template <int SL, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT <= 256 ? 6 : 4))) void k_chem_layer(LayerArgs L) {
  extern __shared__ double smem[];
  __shared__ double ysum3[3];
  __shared__ double dd_s[VK_G * SL];  // A+Ai (+ bottom vdep for M_LHS) per species
  __shared__ double bt_s[VK_G * SL];  // bottom vdep term, kept apart for M_JY
  __shared__ int fix_s[VK_G * SL];    // M_LHS: row i of layer z is fixed
  // Workgroups go to the 8 XCDs round-robin by linear id; the mapping below
  // gives each XCD a contiguous run of (column, layer) work items, so
  // neighbouring layers of a column share one L2 (the y rows of z-1 .. z+1,
  // and the cache lines of a rate table read in its (S, nz) layout).
  int z, c;
  {
    const int64_t W = (int64_t)gridDim.x * gridDim.y;
    const int64_t lin = blockIdx.x + (int64_t)blockIdx.y * gridDim.x;
    const int64_t per = W / 8;
    const int64_t g = (lin < 8 * per) ? (lin % 8) * per + lin / 8 : lin;
    z = (int)(g % gridDim.x);
    c = (int)(g / gridDim.x);
  }
  if (c >= L.ncols) return;
  if (L.active && (L.active[c] & VK_ST_DONE)) return;
  const VkCol col = L.cols[c];
  if (vk_slots(col.ni) != SL) return;  // columns of other buckets
  const int nz = col.nz, ni = col.ni, tid = threadIdx.x;
  if (z >= nz) return;
  const VkNetDev net = L.nets[col.net];
  const int S = net.S, T = net.T;
  const bool want_jac = (L.mode == M_CHEM_JAC || L.mode == M_LHS || L.mode == M_JY);
  const bool want_diff = (L.mode == M_LHS || L.mode == M_RHS2 || L.mode == M_FTOT || L.mode == M_JY);
  const bool want_rhs = (L.mode != M_CHEM_JAC && L.mode != M_JY);
  // rate[S] and drate[S * T] hold 0.0: the record streams point their
  // term-less records there, so every record adds c * (a table entry)
  double* rate = smem;
  double* drate = smem + S + 1;
  double* yrow = smem + S + 1 + (want_jac ? (int64_t)S * T + 1 : 0);  // [3][ni+1]
  if (tid == 0) {
    rate[S] = 0.0;
    if (want_jac) drate[S * T] = 0.0;
  }
  const double* y = L.y + col.off_s;
  const double* atmb = L.atm ? L.atm + col.off_a : nullptr;
  AtmView av;
  if (atmb) av = atm_view(atmb, nz, ni);
  const double Mz = L.M ? L.M[col.off_s / ni + z] : av.M[z];
  // rate of row r at this layer: k[r * nz + z], or kT[z * S + r] (contiguous)
  const double* kz = L.k + col.off_k + (L.k_layer_major ? (int64_t)z * S : (int64_t)z);
  const int64_t kstride = L.k_layer_major ? 1 : nz;

  // y rows z-1, z, z+1 with a trailing 1.0 pad (PAD index = ni)
  for (int e = tid; e < 3 * (ni + 1); e += NT) {
    const int rr = e / (ni + 1), i = e % (ni + 1);
    const int zz = z - 1 + rr;
    double v = 1.0;
    if (i < ni) v = (zz >= 0 && zz < nz) ? y[(int64_t)zz * ni + i] : 0.0;
    yrow[e] = v;
  }
  __syncthreads();
  const double* y0 = yrow + (ni + 1);

  // reaction rates and rate derivatives (chem.py:92-99, 162-183)
  if (T == 2) {
    // Two rows per iteration with every table load issued first; the
    // arithmetic is the general loop's below, term for term.
    auto row2 = [&](int r, double kr, int sp0, int sp1, double st0, double st1, bool tb) {
#pragma clang fp contract(off)
      const double f0 = (st0 > 0.0) ? vk_ipow(y0[sp0], st0) : 1.0;
      const double f1 = (st1 > 0.0) ? vk_ipow(y0[sp1], st1) : 1.0;
      double rt = kr * ((1.0 * f0) * f1);
      if (tb) rt = rt * Mz;
      rate[r] = rt;
      if (want_jac) {
        const double pm0 = (st0 > 0.0) ? vk_ipow(y0[sp0], st0 - 1.0) : 0.0;
        const double pm1 = (st1 > 0.0) ? vk_ipow(y0[sp1], st1 - 1.0) : 0.0;
        double d0 = st0 * pm0 * (1.0 * f1) * kr;
        double d1 = st1 * pm1 * (1.0 * f0) * kr;
        if (tb) {
          d0 = d0 * Mz;
          d1 = d1 * Mz;
        }
        drate[2 * r] = d0;
        drate[2 * r + 1] = d1;
      }
    };
    for (int r0 = tid; r0 < S; r0 += 2 * NT) {
      const int r1 = r0 + NT;
      const bool has1 = r1 < S;
      const int rb = has1 ? r1 : r0;
      const double ka = kz[(int64_t)r0 * kstride], kb = kz[(int64_t)rb * kstride];
      const int2 ia = *reinterpret_cast<const int2*>(net.r_idx + 2 * r0);
      const int2 ib = *reinterpret_cast<const int2*>(net.r_idx + 2 * rb);
      const double2 sa = *reinterpret_cast<const double2*>(net.r_st + 2 * r0);
      const double2 sb = *reinterpret_cast<const double2*>(net.r_st + 2 * rb);
      const int ta = net.is3[r0], tbb = net.is3[rb];
      row2(r0, ka, ia.x, ia.y, sa.x, sa.y, ta != 0);
      if (has1) row2(r1, kb, ib.x, ib.y, sb.x, sb.y, tbb != 0);
    }
  } else
  for (int r = tid; r < S; r += NT) {
#pragma clang fp contract(off)
    const double kr = kz[(int64_t)r * kstride];
    double fac[4];
    double prodr = 1.0;
    for (int s = 0; s < T; ++s) {
      const int sp = net.r_idx[r * T + s];
      const double st = net.r_st[r * T + s];
      const double f = (st > 0.0) ? vk_ipow(y0[sp], st) : 1.0;
      if (s < 4) fac[s] = f;
      prodr = prodr * f;
    }
    double rt = kr * prodr;
    const bool tb = net.is3[r] != 0;
    if (tb) rt = rt * Mz;
    rate[r] = rt;
    if (want_jac) {
      for (int s = 0; s < T; ++s) {
        const int sp = net.r_idx[r * T + s];
        const double st = net.r_st[r * T + s];
        double leave = 1.0;
        for (int l2 = 0; l2 < T; ++l2)
          if (l2 != s) leave = leave * (l2 < 4 ? fac[l2] : 1.0);
        const double pm1 = (st > 0.0) ? vk_ipow(y0[sp], st - 1.0) : 0.0;
        double dr = st * pm1 * leave * kr;
        if (tb) dr = dr * Mz;
        drate[r * T + s] = dr;
      }
    }
  }
  // column sums of y for the eddy terms (gas species only)
  if (want_diff && (tid & 63) == 0 && (tid >> 6) < 3) {
    // row rr = wave: a sequential sum in one lane keeps it reproducible;
    // non-gas species add +0.0, which leaves the sum unchanged
    const int rr = tid >> 6, zz = z - 1 + rr;
    double v = 1.0;
    if (zz >= 0 && zz < nz) {
      v = 0.0;
      for (int i = 0; i < ni; ++i) v += (av.gas[i] != 0.0) ? yrow[rr * (ni + 1) + i] : 0.0;
    }
    ysum3[rr] = v;
  }
  __syncthreads();
  // loss / production chunk sums from the lane-balanced records
  double* part = yrow + 3 * (ni + 1);
  if (want_rhs) {
    for (int vl = tid; vl < VK_CHEM_NT; vl += NT) {
      double acc = 0.0;
      // a term-less record adds c * rate[S] = +-0, which leaves acc (never
      // -0: it starts at +0) unchanged
      walk_records<true>(
          net.frec + vl, net.frec_k,
          [&](const VkRec& rc) { return (double)rc.c * rate[vk_rec_item(rc.w)]; },
          [&](const VkRec& rc, double t) {
#pragma clang fp contract(off)
            acc = acc + t;
            const uint32_t dest = vk_rec_dest(rc.w);
            if (dest != VK_REC_NONE_DEST) {
              part[dest] = acc;
              acc = 0.0;
            }
          });
    }
    __syncthreads();
  }

  const double ysm = want_diff ? ysum3[0] : 0.0;
  const double ys0 = want_diff ? ysum3[1] : 0.0;
  const double ysp = want_diff ? ysum3[2] : 0.0;
  DiffCoef ed = {0.0, 0.0, 0.0};
  if (want_diff) ed = eddy_coef(z, nz, av.Kzz, av.dzi, av.vz, ysm, ys0, ysp);
  const bool use_top = want_diff && av.flags[2] != 0.0;
  const bool use_bot = want_diff && av.flags[3] != 0.0;
  const double dtc = L.dt ? L.dt[c] : 1.0;
  const double c0 = 1.0 / (GAMMA * dtc);

  // per-species outputs: chem rhs (+ diffusion), diagonal/off-diagonal terms
  for (int i = tid; i < ni; i += NT) {
#pragma clang fp contract(off)
    double f = 0.0;
    if (want_rhs) {
      double loss = 0.0, prod = 0.0;
      for (int ch = net.lc_ptr[i]; ch < net.lc_ptr[i + 1]; ++ch) loss = loss + part[ch];
      for (int ch = net.pc_ptr[i]; ch < net.pc_ptr[i + 1]; ++ch) prod = prod + part[ch];
      f = prod - loss;
    }
    const int64_t e_out = col.off_s + (int64_t)z * ni + i;
    if (want_diff) {
      const DiffCoef md =
          L.mstat ? mol_coef_pre(z, nz, L.mstat + col.off_s + (int64_t)z * ni + i, L.mstride, ysm,
                                 ys0, ysp)
                  : mol_coef(z, nz, ni, i, av, ysm, ys0, ysp);
      const double At = ed.A + md.A, Bt = ed.B + md.B, Ct = ed.C + md.C;
      const double yi = y0[i];
      double diff;
      if (z == 0)
        diff = At * yi + Bt * yrow[2 * (ni + 1) + i];
      else if (z == nz - 1)
        diff = At * yi + Ct * yrow[i];
      else
        diff = At * yi + Bt * yrow[2 * (ni + 1) + i] + Ct * yrow[i];
      if (z == nz - 1 && use_top) diff = diff + av.top_flux[i] / av.dzi[nz - 2];
      if (z == 0 && use_bot) diff = diff + (av.bot_flux[i] - yi * av.bot_vdep[i]) / av.dzi[0];
      const double bterm = (z == 0 && use_bot) ? (-av.bot_vdep[i] / av.dzi[0]) : 0.0;
      if (want_jac) {
        dd_s[i] = (L.mode == M_LHS && z == 0 && use_bot) ? At + bterm : At;
        bt_s[i] = bterm;
      }
      const bool fixed = (L.fix != nullptr) && L.fix[e_out];
      if (L.mode == M_LHS) {
        fix_s[i] = fixed ? 1 : 0;
        const double tot = f + diff;
        L.out0[e_out] = fixed ? 0.0 : tot;
        if (z < nz - 1) {
          const bool fx = fixed;
          L.out_sup[col.off_s + (int64_t)z * ni + i] = fx ? 0.0 : -Bt;
        }
        if (z > 0) L.out_sub[col.off_s + (int64_t)(z - 1) * ni + i] = fixed ? 0.0 : -Ct;
      } else if (L.mode == M_JY) {
        if (z < nz - 1) L.out_sup[col.off_s + (int64_t)z * ni + i] = Bt;
        if (z > 0) L.out_sub[col.off_s + (int64_t)(z - 1) * ni + i] = Ct;
      } else if (L.mode == M_RHS2) {
        const double tot = fixed ? 0.0 : (f + diff);
        L.out0[e_out] = tot - (2.0 / (GAMMA * dtc)) * L.k1[e_out];
      } else {  // M_FTOT
        L.out0[e_out] = f + diff;
      }
    } else if (L.mode == M_CHEM_RHS) {
      L.out0[e_out] = f;
    }
  }
  if (!want_jac) return;
  __syncthreads();
  // Jacobian entries from the entry-major CSR (chem.py:187-208)
  if (L.mode == M_LHS) {
    // tile-layout D = c0 I - J - diag(dd), fix rows -> c0 e_i.  Each tile
    // entry is closed by one record of one lane (the jac list of (i, q) in
    // order), which writes 0 - J to its tile slot (padding entries have no
    // terms: +0); then thread i finishes row i's diagonal entry, (c0 + (0 - J))
    // + (-dd_i), the same bits as (c0 - J) + (-dd_i), or writes the whole row
    // of a fixed species.
    constexpr int BLK = vk_blk<SL>();
    double* dst = L.out_blk + col.off_b + (int64_t)z * BLK;
    for (int vl = tid; vl < VK_CHEM_NT; vl += NT) {
      double J = 0.0;
      walk_records<false>(
          net.jrec + vl, net.jrec_k,
          [&](const VkRec& rc) { return (double)rc.c * drate[vk_rec_item(rc.w)]; },
          [&](const VkRec& rc, double t) {
#pragma clang fp contract(off)
            J = J + t;  // term-less records add +-0 (see rate[S])
            const uint32_t dest = vk_rec_dest(rc.w);
            if (dest != VK_REC_NONE_DEST) {
              dst[dest] = 0.0 - J;
              J = 0.0;
            }
          });
    }
    __syncthreads();
    for (int i = tid; i < ni; i += NT) {
#pragma clang fp contract(off)
      const int a = i / VK_G, r = i % VK_G;
      if (fix_s[i]) {
        for (int q = 0; q < ni; ++q)
          dst[((a * SL + q / VK_G) * VK_G + r) * VK_G + q % VK_G] = (q == i) ? c0 : 0.0;
      } else {
        double* e = dst + ((a * SL + a) * VK_G + r) * VK_G + r;
        *e = (c0 + *e) + (-dd_s[i]);
      }
    }
  } else {
    double* dst = L.out0 + col.off_s * ni + (int64_t)z * ni * ni;
    for (int e = tid; e < ni * ni; e += NT) {
#pragma clang fp contract(off)
      double J = 0.0;
      for (int p = net.jac_ptr[e]; p < net.jac_ptr[e + 1]; ++p)
        J = J + net.jac_c[p] * drate[net.jac_item[p]];
      if (L.mode == M_JY) {
        const int i = e / ni, q = e % ni;
        if (i == q) {
          J = J + dd_s[i];
          if (z == 0 && use_bot) J = J + bt_s[i];
        }
      }
      dst[e] = J;
    }
  }
}

// Largest dynamic LDS a launch of fn may use; raises the function's limit to it.
static int raise_lds_limit(const void* fn) {
  int dev = 0, per_block = 64 * 1024;
  hipFuncAttributes fa{};
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
          hipSuccess ||
      hipFuncGetAttributes(&fa, fn) != hipSuccess) {
    (void)hipGetLastError();
    return 64 * 1024 - 2048;
  }
  const int lim = per_block - (int)fa.sharedSizeBytes;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) {
    (void)hipGetLastError();
    return 64 * 1024 - (int)fa.sharedSizeBytes;
  }
  return lim;
}

int vk_launch_chem_layer(int sl, int mode, const VkCol* cols, int ncols, int max_nz, int max_S,
                         int max_T, const VkNetDev* nets, const double* y, const double* M,
                         const double* atm, const double* k, const double* dt, const uint8_t* fix,
                         const double* k1, double* out0, double* out_blk, double* out_sup,
                         double* out_sub, const int* active, int k_layer_major,
                         const double* mstat, int64_t mstride, hipStream_t stream) {
  LayerArgs L{cols, ncols, nets, y, M, atm, k, dt, fix, k1, out0, out_blk, out_sup, out_sub,
              active, mode, k_layer_major, mstat, mstride};
  const bool want_jac = (mode == M_CHEM_JAC || mode == M_LHS || mode == M_JY);
  // ni+1 <= 16*sl+1 for the y rows; loss / production chunks: at most
  // 2*S*T list entries in at most 2*ni lists
  const size_t n16 = (size_t)VK_G * (sl > 0 ? sl : 1);
  const size_t nchunk_max = (2 * (size_t)max_S * max_T + VK_FCHUNK - 1) / VK_FCHUNK + 2 * n16;
  const size_t shmem =
      sizeof(double) * ((size_t)max_S + 1 + (want_jac ? (size_t)max_S * max_T + 1 : 0) +
                        3 * (n16 + 1) + nchunk_max);
  if (shmem > 160 * 1024) {
    vk_set_error("chem layer needs %zu bytes of LDS (S=%d, T=%d)", shmem, max_S, max_T);
    return VK_ERR_UNSUPPORTED;
  }
  dim3 grid(max_nz, ncols);
  // The dynamic-LDS limit of each instantiation is raised once (to what the
  // device allows beyond the kernel's static LDS), outside any graph capture
  // that later launches may be part of.
  static int limit[2][8] = {{0}};
  const int big = want_jac ? 1 : 0;
  switch (sl) {
#define VK_CASE(n)                                                                        \
  case n: {                                                                               \
    const void* fn = big ? (const void*)k_chem_layer<n, VK_CHEM_NT>                       \
                         : (const void*)k_chem_layer<n, VK_CHEM_NT / 2>;                  \
    if (!limit[big][n]) limit[big][n] = raise_lds_limit(fn);                              \
    if ((int)shmem > limit[big][n]) {                                                     \
      vk_set_error("chem layer needs %zu bytes of dynamic LDS, limit %d", shmem,         \
                   limit[big][n]);                                                        \
      return VK_ERR_UNSUPPORTED;                                                          \
    }                                                                                     \
    if (big)                                                                              \
      hipLaunchKernelGGL((k_chem_layer<n, VK_CHEM_NT>), grid, dim3(VK_CHEM_NT), shmem,    \
                         stream, L);                                                      \
    else                                                                                  \
      hipLaunchKernelGGL((k_chem_layer<n, VK_CHEM_NT / 2>), grid, dim3(VK_CHEM_NT / 2),  \
                         shmem, stream, L);                                               \
  } break;
    VK_CASE(1) VK_CASE(2) VK_CASE(3) VK_CASE(4) VK_CASE(5) VK_CASE(6) VK_CASE(7)
#undef VK_CASE
    default:
      return VK_ERR_UNSUPPORTED;
  }
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

// kT[off_k + z*S + r] = k[off_k + r*nz + z]: the rate tables layer-major,
// through an LDS tile of 64 rows x 32 layers so both sides are coalesced
// (256-byte runs of layers read, 512-byte runs of rows written); each thread
// moves eight elements.
constexpr int KT_R = 64, KT_Z = 32;
__global__ __launch_bounds__(256) void k_rates_layer_major(const VkCol* __restrict__ cols, int ncols,
                                                          const double* __restrict__ k,
                                                          double* __restrict__ kT,
                                                          const int* __restrict__ active, int c0) {
  __shared__ double tile[KT_R][KT_Z + 1];
  const int c = c0 + (int)blockIdx.z;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  const int S = col.S, nz = col.nz;
  const int r0 = blockIdx.x * KT_R, z0 = blockIdx.y * KT_Z;
  if (r0 >= S || z0 >= nz) return;
  const int tid = threadIdx.x;
  {
    const int zl = tid & (KT_Z - 1), rl = tid >> 5;  // 32 layers x 8 rows
    const int z = z0 + zl;
#pragma unroll
    for (int i = rl; i < KT_R; i += 8) {
      const int r = r0 + i;
      if (r < S && z < nz) tile[i][zl] = k[col.off_k + (int64_t)r * nz + z];
    }
  }
  __syncthreads();
  {
    const int rl = tid & (KT_R - 1), zl0 = tid >> 6;  // 64 rows x 4 layers
    const int r = r0 + rl;
#pragma unroll
    for (int i = zl0; i < KT_Z; i += 4) {
      const int z = z0 + i;
      if (r < S && z < nz) kT[col.off_k + (int64_t)z * S + r] = tile[rl][i];
    }
  }
}

int vk_launch_rates_layer_major(const VkCol* cols, int ncols, int max_S, int max_nz, const double* k,
                                double* kT, const int* active, hipStream_t stream) {
  for (int c0 = 0; c0 < ncols; c0 += 65535) {  // grid z is limited to 65535
    const int nc = ncols - c0 < 65535 ? ncols - c0 : 65535;
    dim3 grid((max_S + KT_R - 1) / KT_R, (max_nz + KT_Z - 1) / KT_Z, nc);
    hipLaunchKernelGGL(k_rates_layer_major, grid, dim3(256), 0, stream, cols, ncols, k, kT, active,
                       c0);
    VK_CHECK_HIP(hipGetLastError());
  }
  return VK_OK;
}

// ms[c * stride + off_s + z*ni + i] = component c of mol_static for species
// i at layer z of column c (grid: layers x columns, one thread per species).
__global__ __launch_bounds__(128) void k_mol_static(const VkCol* __restrict__ cols, int ncols,
                                                    const double* __restrict__ atm,
                                                    double* __restrict__ ms, int64_t stride,
                                                    const int* __restrict__ active) {
  const int c = blockIdx.y, z = blockIdx.x;
  if (c >= ncols) return;
  if (active && (active[c] & VK_ST_DONE)) return;
  const VkCol col = cols[c];
  if (z >= col.nz) return;
  const AtmView av = atm_view(atm + col.off_a, col.nz, col.ni);
  for (int i = threadIdx.x; i < col.ni; i += blockDim.x) {
    double m[8];
    mol_static(z, col.nz, col.ni, i, av, m);
    double* o = ms + col.off_s + (int64_t)z * col.ni + i;
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q * stride] = m[q];
  }
}

// compute_diff_grav (jax_step.py:67-147): the y-independent gravity /
// thermal-diffusion, molecular-advection and settling parts of the
// molecular-diffusion coefficients, the additive terms of mol_static.
// A, B, C are (nz, ni) per column at off_s: rows 1..nz-2 hold A/B/C_grav_int,
// row 0 holds bdry0_grav (A, B), row nz-1 bdry_top_grav (A, C); the entries
// the reference does not define are zero.
__global__ __launch_bounds__(128) void k_diff_grav(const VkCol* __restrict__ cols, int ncols,
                                                   const double* __restrict__ atm,
                                                   double* __restrict__ A, double* __restrict__ B,
                                                   double* __restrict__ C) {
  const int c = blockIdx.y, z = blockIdx.x;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  if (z >= col.nz) return;
  const AtmView av = atm_view(atm + col.off_a, col.nz, col.ni);
  for (int i = threadIdx.x; i < col.ni; i += blockDim.x) {
    double m[8];
    mol_static(z, col.nz, col.ni, i, av, m);
    const int64_t e = col.off_s + (int64_t)z * col.ni + i;
    A[e] = m[5];
    B[e] = m[6];
    C[e] = m[7];
  }
}

int vk_launch_diff_grav(const VkCol* cols, int ncols, int max_nz, const double* atm, double* A,
                        double* B, double* C, hipStream_t stream) {
  if (ncols <= 0 || max_nz <= 0) return VK_OK;
  hipLaunchKernelGGL(k_diff_grav, dim3(max_nz, ncols), dim3(128), 0, stream, cols, ncols, atm, A,
                     B, C);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_mol_static(const VkCol* cols, int ncols, int max_nz, const double* atm, double* ms,
                         int64_t stride, const int* active, hipStream_t stream) {
  if (ncols <= 0 || max_nz <= 0) return VK_OK;
  hipLaunchKernelGGL(k_mol_static, dim3(max_nz, ncols), dim3(128), 0, stream, cols, ncols, atm,
                     ms, stride, active);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
