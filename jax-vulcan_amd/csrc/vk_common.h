// Shared definitions for the gfx950 Ros2 column kernels.
//
// Layout conventions (all FP64 unless noted, row-major):
//   * a "column" is one atmosphere: nz layers x ni species;
//   * per-column arrays live in packed batch buffers at offsets given by a
//     VkCol descriptor, so one launch can carry columns of different nz;
//   * dense ni x ni blocks (Jacobian blocks, block inverses) are stored in
//     the "tile layout" of the 16x16 thread grid that factors them: element
//     (i, j) with i = 16a + tr, j = 16b + tc lives at ((a*SL + b)*256 +
//     tr*16 + tc), so every load/store of a register tile is a contiguous
//     512-byte wave access.  SL = ceil(ni/16) is a template parameter.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vulcan_ros2.h"

#define VK_NT 256          // threads per column workgroup (16 x 16 grid)
#define VK_G 16            // grid edge

// Status bits written per column (never a batch abort).
#define VK_ST_SINGULAR 1   // zero pivot met in a block inverse
#define VK_ST_NONFINITE 2  // non-finite value in the state after a step
#define VK_ST_CONVERGED 4  // steady-state criterion met
#define VK_ST_MAXSTEPS 8   // step cap reached
#define VK_ST_DONE (VK_ST_CONVERGED | VK_ST_MAXSTEPS)

// Offsets (in doubles) of the fields of one packed atmosphere record.
struct VkAtmOff {
  int64_t Kzz, dzi, vz, Hpi, Ti, Tco, g, M;
  int64_t ms, alpha, top_flux, bot_flux, bot_vdep, gas;
  int64_t Dzz, vm, vs, flags, total;
};

__host__ __device__ inline VkAtmOff vk_atm_offsets(int nz, int ni) {
  VkAtmOff o;
  int64_t p = 0;
  const int64_t nzm = nz - 1;
  o.Kzz = p; p += nzm;
  o.dzi = p; p += nzm;
  o.vz = p; p += nzm;
  o.Hpi = p; p += nzm;
  o.Ti = p; p += nzm;
  o.Tco = p; p += nz;
  o.g = p; p += nz;
  o.M = p; p += nz;
  o.ms = p; p += ni;
  o.alpha = p; p += ni;
  o.top_flux = p; p += ni;
  o.bot_flux = p; p += ni;
  o.bot_vdep = p; p += ni;
  o.gas = p; p += ni;
  o.Dzz = p; p += nzm * ni;
  o.vm = p; p += (int64_t)nz * ni;
  o.vs = p; p += nzm * ni;
  o.flags = p; p += 4;
  o.total = p;
  return o;
}

// Device copy of one reaction network.  Index tables are int32; the
// stoichiometric coefficients stay float64 exactly as in the reference
// tables.  Three CSR structures give deterministic, atomic-free sums:
//   loss/prod  : species-major lists of (rate row*T + slot) in ascending
//                (row, slot) order, the order np.add.at / a sequential
//                segment_sum visit them;
//   jac        : entry-major lists over (out, in) pairs of
//                (row*T + in_slot, signed out stoich) in ascending
//                (row, out slot, in slot) order.
// Threads per workgroup of the chemistry kernels (one (layer, column) each).
constexpr int VK_CHEM_NT = 512;

// Lane-balanced record streams, built once per network from the lists
// above, for the VK_CHEM_NT lanes of a chemistry workgroup.  Record k of
// lane l sits at [k * VK_CHEM_NT + l], so a wave's loads are
// contiguous, and lanes carry near-equal record counts however uneven the
// lists are.  A record adds c * src[item] to the lane's running sum (a
// record with no term points item at the zero slot after the table: rate[S]
// for frec, drate[S * T] for jrec); a record whose dest is not
// VK_REC_NONE_DEST closes the sum into output dest and restarts it at zero.
//   jrec: one output per tile entry of D (dest = tile index), whose terms
//         are its jac list in order, so each entry is the same sequential
//         sum as a walk of that list (src = drate);
//   frec: one output per chunk of at most VK_FCHUNK consecutive entries of a
//         loss or production list (dest = chunk, src = rate, item = row);
//         lc_ptr / pc_ptr give species i's loss / production chunks, summed
//         in list order.
// Records are 8 bytes: the coefficient as a float (network creation checks
// that every stoichiometric coefficient is exact in float, so widening it
// back gives the float64 value) and item | dest << 18 in one word.
struct VkRec {
  float c;
  uint32_t w;
};
constexpr uint32_t VK_REC_NONE_ITEM = 0x3FFFFu;  // 18 bits: S * T must stay below
constexpr uint32_t VK_REC_NONE_DEST = 0x3FFFu;   // 14 bits
__host__ __device__ constexpr uint32_t vk_rec_item(uint32_t w) { return w & 0x3FFFFu; }
__host__ __device__ constexpr uint32_t vk_rec_dest(uint32_t w) { return w >> 18; }
__host__ __device__ constexpr uint32_t vk_rec_word(uint32_t item, uint32_t dest) {
  return item | (dest << 18);
}
constexpr int VK_FCHUNK = 16;
// Records per lane are a multiple of VK_REC_GROUP: the kernels load a group
// of records at once and the next group before working through the current.
constexpr int VK_REC_GROUP = 8;

struct VkNetDev {
  int ni, S, T, pad;
  const VkRec* jrec;     // (jrec_k * VK_NT)
  const VkRec* frec;     // (frec_k * VK_NT)
  const int* lc_ptr;     // (ni+1) loss chunks of species i
  const int* pc_ptr;     // (ni+1) production chunks of species i
  int jrec_k, frec_k, nchunk, pad2;
  const int* r_idx;      // (S*T) reactant species, PAD = ni
  const double* r_st;    // (S*T)
  const int* p_idx;      // (S*T) product species, PAD = ni
  const double* p_st;    // (S*T)
  const int* is3;        // (S) three-body flag
  const int* loss_ptr;   // (ni+1)
  const int* loss_item;  // row*T + slot
  const double* loss_c;  // stoich
  const int* prod_ptr;
  const int* prod_item;
  const double* prod_c;
  const int* jac_ptr;    // (ni*ni+1), entry e = out*ni + in
  const int* jac_item;   // row*T + in_slot
  const double* jac_c;   // signed out stoich
};

// Everything the rate build reads of one (forward, reverse) row pair: row
// i = 2p + 1 and its reverse i + 1, in one 128-byte record so a wave fetches
// a pair's tables in one go (vk_rates_create fills it).
enum : int {
  VK_PAIR_ZERO_FWD = 1,  // forward row held at zero in the full build
  VK_PAIR_ZERO_REV = 2,  // reverse row held at zero (or absent) in the full build
  VK_PAIR_HAS_REV = 4,   // K_eq is computed (else the sentinel 1)
};
struct alignas(16) VkRatePair {
  double par[6];      // a, n, E, a_inf, n_inf, E_inf
  double dn;          // sum of reactant minus sum of product stoichiometries
  float st[8];        // reactant stoich of slots 0..3, product stoich at 4..7
  uint8_t sp[8];      // species of those slots, PAD = ni
  int32_t form;       // VK_KF_*
  int32_t lowt;       // VK_LOWT_*
  int32_t flags;      // VK_PAIR_*
  int32_t pad[5];
};
static_assert(sizeof(VkRatePair) == 128, "VkRatePair is one 128-byte line");

// Device copy of one network's rate-construction tables (vk_rates_create).
struct VkRatesDev {
  int nr, ni, npairs, stop_rev;
  const VkRatePair* pairs;  // (npairs)
  const double* nasa9;      // (ni, 2, 10) or null
  const uint8_t* zero_rows; // (nr+1) or null
};

// Refresh-record offsets of one column (vk_refresh_record_layout).
struct VkRefOff {
  int64_t pico, mass, esc, scal, total;
};
__host__ __device__ inline VkRefOff vk_refresh_offsets(int nz, int ni) {
  VkRefOff o;
  o.pico = 0;
  o.mass = nz + 1;
  o.esc = o.mass + ni;
  o.scal = o.esc + ni;
  o.total = o.scal + 7;
  return o;
}

// The controller's clip of a step output (outer_loop.py:196-224).
__device__ inline double vk_clip_val(double v, double ymix_old, const vk_ss_params& p) {
  double c = (v < p.pos_cut && v >= p.nega_cut) ? 0.0 : v;
  if (ymix_old < p.mtol && c < 0.0) c = 0.0;
  return c;
}

// One column of a batch.
struct VkCol {
  int32_t nz, ni, S, net;
  int64_t off_s;  // (nz, ni) state-shaped arrays
  int64_t off_k;  // (S, nz) rate table
  int64_t off_a;  // packed atmosphere record
  int64_t off_b;  // (nz, SL*SL*256) block stacks
};

template <int SL>
__host__ __device__ constexpr int vk_blk() { return SL * SL * VK_NT; }

__host__ __device__ inline int vk_slots(int ni) { return (ni + VK_G - 1) / VK_G; }

// Epilogue applied by the backward sweep to each element of x_j (all
// pointers at off_s):
//   kind 1: out1 = y + x / e0                         (Ros2 stage-1 yk2)
//   kind 2: out1 = y + e0*k1 + e1*x, out2 = |out1 - yk2|  (Ros2 sol, delta)
// block-factor kernel variant per SL bucket (VK_FACTOR_* in vulcan_ros2.h)
extern int g_factor_variant[8];

struct SolveEpi {
  int kind;
  double e0, e1;
  const double* y;
  const double* k1;
  const double* yk2;
  double* out1;
  double* out2;
  // when set: the solution is base + the solve's result, i.e. the solve
  // computed a refinement correction; base may be the output x itself
  // (each element is read, then written, by one thread)
  const double* base = nullptr;
};

// Thread-local error text set by the host API.
void vk_set_error(const char* fmt, ...);

#define VK_CHECK_HIP(expr)                                                    \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      vk_set_error("%s failed: %s", #expr, hipGetErrorString(_e));           \
      return VK_ERR_HIP;                                                      \
    }                                                                         \
  } while (0)

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

// v plus v rotated right by N lanes within each 16-lane row (DPP row_ror).
template <int N>
__device__ __forceinline__ double vk_row_ror_add(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x120 + N, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x120 + N, 0xF, 0xF, false);
  return v + __hiloint2double(hi, lo);
}

// Sum over the 16 lanes that share a grid row (lanes tr*16 .. tr*16+15), in
// every lane.  After the rotation by 8 a lane's value depends on its index
// mod 8 only, so rotating by 4, 2, 1 next pairs lane i with lane i^4, i^2,
// i^1: the butterfly sum ((v_i + v_i^8) + ...), the same in every lane.
__device__ inline double vk_rowsum16(double v) {
  v = vk_row_ror_add<8>(v);
  v = vk_row_ror_add<4>(v);
  v = vk_row_ror_add<2>(v);
  v = vk_row_ror_add<1>(v);
  return v;
}

// Lane KC of every 16-lane row, broadcast to the whole row (DPP row_newbcast).
// Every lane is written, so the DPP move needs no old value; on the 64-bit
// value this is one v_mov_b64 with DPP (row_newbcast is a DPP64 control).
template <int KC>
__device__ __forceinline__ double row_bcast(double v) {
  static_assert(sizeof(long) == sizeof(double), "64-bit long");
  const long x = __builtin_bit_cast(long, v);
  const long y = __builtin_amdgcn_mov_dpp(x, 0x150 + KC, 0xF, 0xF, false);
  return __builtin_bit_cast(double, y);
}

// v = max(v, v moved by a DPP pattern); rows outside ROW_MASK keep v.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_fmax(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int lo2 = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
  const int hi2 = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
  return fmax(v, __hiloint2double(hi2, lo2));
}

// Max over a wave whose 16-lane rows each hold one value in every lane (no
// NaN): row_bcast 15 and 31 carry rows 0-2 into row 3, whose lane 63 is read
// back as a scalar.
__device__ __forceinline__ double rows_max(double v) {
  v = dpp_fmax<0x142, 0xA>(v);
  v = dpp_fmax<0x143, 0xC>(v);
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// Integer power for small non-negative integer exponents held as double.
__device__ inline double vk_ipow(double y, double nu) {
  int n = (int)nu;
  double r = y;
  if (n == 1) return y;
  if (n == 2) return y * y;
  if (n <= 0) return 1.0;
  for (int i = 1; i < n; ++i) r = r * y;
  return r;
}
