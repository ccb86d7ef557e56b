// Host side of the C ABI: error text, network upload, batch tables and the
// entry points that sequence the kernels on the caller's stream.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <numeric>
#include <queue>
#include <set>
#include <vector>

#include "vk_common.h"

namespace {
thread_local char g_err[512] = "";
}

void vk_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------------------
// optional per-kernel timing with HIP events recorded on the launch stream
// ---------------------------------------------------------------------------
namespace {
enum { VK_NKERN = 7 };
struct TimedLaunch {
  int kind;
  hipEvent_t a, b;
};
struct Timing {
  bool on = false;
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> pool;
  double ms[VK_NKERN] = {0};
  long long count[VK_NKERN] = {0};
};
Timing g_timing;
// Set while the allocation-free step runs: no timing events are taken.
thread_local bool g_timing_suppressed = false;

hipEvent_t take_event() {
  if (!g_timing.pool.empty()) {
    hipEvent_t e = g_timing.pool.back();
    g_timing.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct KernelTimer {
  int kind;
  hipStream_t st;
  hipEvent_t a = nullptr, b = nullptr;
  KernelTimer(int k, hipStream_t s) : kind(k), st(s) {
    if (!g_timing.on || g_timing_suppressed) return;
    a = take_event();
    b = take_event();
    if (a && b) (void)hipEventRecord(a, st);
  }
  ~KernelTimer() {
    if (!a || !b) return;
    (void)hipEventRecord(b, st);
    g_timing.pending.push_back({kind, a, b});
  }
};
}  // namespace

int vk_launch_bt_factor(int sl, const VkCol* cols, int ncols, const double* diag, int diag_tile,
                        int transpose, const double* sup_d, const double* sub_d, double* inv,
                        const double* rhs, double* rmod, int* status, const int* active,
                        hipStream_t stream, double* schur = nullptr);
bool vk_factor_fuses_sweep(int sl);
bool vk_factor_stores_schur(int sl);
extern int g_solve_hold;
bool vk_factor_is_cr(int sl);
int vk_launch_nat_to_tile(int sl, const VkCol* cols, int ncols, int max_nz, const double* nat,
                          double* tile, int transpose, hipStream_t stream);
int vk_launch_cr_factor(int sl, const VkCol* cols, int ncols, int max_nz, const double* blk,
                        const double* sup_d, const double* sub_d, double* cr, int* status,
                        const int* active, hipStream_t stream, const double* rhs_fused = nullptr);
int vk_launch_cr_solve(int sl, const VkCol* cols, int ncols, int max_nz, const double* rhs,
                       const double* sup_d, const double* sub_d, double* cr, double* x,
                       SolveEpi epi, const int* active, hipStream_t stream,
                       bool fwd_done = false);
int vk_launch_bt_solve(int sl, const VkCol* cols, int ncols, const double* inv,
                       const double* sup_d, const double* sub_d, const double* rhs,
                       double* rscratch, double* x, int mode, SolveEpi epi, const int* active,
                       hipStream_t stream, const double* refine_diag = nullptr,
                       int* refl = nullptr, const double* refine_schur = nullptr);
int vk_launch_tile_to_nat(int sl, const VkCol* cols, int ncols, int max_nz, const double* tile,
                          double* nat, hipStream_t stream);
int vk_launch_diff_grav(const VkCol* cols, int ncols, int max_nz, const double* atm, double* A,
                        double* B, double* C, hipStream_t stream);
int vk_launch_bt_dense(int sl, const VkCol* cols, int ncols, const double* diag,
                       const double* sup, const double* sub, const double* rhs, double* x,
                       double* inv, double* rmod, int* status, hipStream_t stream);
int vk_launch_bt_residual(int sl, const VkCol* cols, int ncols, int max_nz, const double* blk,
                          const double* sup_d, const double* sub_d, const double* rhs,
                          const double* x, double* res, const int* active, hipStream_t stream);
int vk_launch_chem_layer(int sl, int mode, const VkCol* cols, int ncols, int max_nz, int max_S,
                         int max_T, const VkNetDev* nets, const double* y, const double* M,
                         const double* atm, const double* k, const double* dt, const uint8_t* fix,
                         const double* k1, double* out0, double* out_blk, double* out_sup,
                         double* out_sub, const int* active, int k_layer_major,
                         const double* mstat, int64_t mstride, hipStream_t stream);
int vk_launch_mol_static(const VkCol* cols, int ncols, int max_nz, const double* atm, double* ms,
                         int64_t stride, const int* active, hipStream_t stream);
int vk_launch_rates_layer_major(const VkCol* cols, int ncols, int max_S, int max_nz, const double* k,
                                double* kT, const int* active, hipStream_t stream);

int vk_launch_ss_control(const VkCol* cols, int ncols, const double* atm, const vk_ss_params& p,
                         const vk_ss_state& s, hipStream_t st);
int vk_launch_ss_init_range(const VkCol* cols, int ncols, int first, int count,
                            const double* atm, const vk_ss_params& p, const vk_ss_state& s,
                            hipStream_t st);
int vk_launch_ss_init(const VkCol* cols, int ncols, const double* atm, const vk_ss_params& p,
                      const vk_ss_state& s, hipStream_t st);

int vk_launch_photo(int parts, const VkCol* cols, int ncols, int max_nz,
                    const vk_photo_setup& P, const vk_photo_state& S, const double* y, double* k,
                    hipStream_t st);

int vk_launch_delta_max(const VkCol* cols, int ncols, const double* sol, const double* delta,
                        double* out, int64_t stride, hipStream_t st);
int vk_launch_ss_vjp(const VkCol* cols, int ncols, int max_nz, int max_ni, int64_t max_rec,
                     const VkNetDev* nets, const double* y, const double* atm, const double* k,
                     const double* lam, double* cot_k, double* cot_atm, hipStream_t st);

struct vk_network_s {
  VkNetDev dev;      // device pointers into buf
  void* buf = nullptr;
  int ni = 0, nr = 0, T = 0;
};

struct vk_batch_s {
  int ncols = 0;
  VkCol* d_cols = nullptr;
  VkNetDev* d_nets = nullptr;
  std::vector<VkCol> cols;
  std::vector<int> slots;   // distinct SL buckets
  int max_nz = 0, max_S = 0, max_T = 0, max_ni = 0;
  int64_t max_rec = 0;
  int64_t state_doubles = 0, block_doubles = 0, k_doubles = 0;
  bool uniform = false;
  // slot batch (continuous batching): per-slot capacity and network table
  bool slots_mode = false;
  int cap_nz = 0, cap_ni = 0, cap_S = 0, cap_T = 0;
  // one helper stream per SL bucket (default priority) and fork/join events:
  // the buckets of an uneven batch run concurrently
  std::vector<hipStream_t> aux;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  // networks: (ni, nr) per batch network, and the rates tables bound to
  // them by vk_batch_set_rates
  std::vector<std::pair<int, int>> net_dims;
  VkRatesDev* d_rates = nullptr;
  int n_rates = 0;
  // latency mode: every bucket of this batch uses the cyclic-reduction
  // factor, whatever the process-wide VK_TUNE_FACTOR says
  bool latency_mode = false;
};

struct vk_rates_s {
  VkRatesDev dev{};
  void* buf = nullptr;
};

namespace {

enum { M_CHEM_RHS = 1, M_CHEM_JAC = 2, M_LHS = 3, M_RHS2 = 4, M_FTOT = 5, M_JY = 6 };

template <class T>
size_t align_up(size_t n) { return (n * sizeof(T) + 255) & ~size_t(255); }

int check_batch(vk_batch b) {
  if (!b) { vk_set_error("null batch"); return VK_ERR_ARG; }
  return VK_OK;
}

// Entry points that touch every column: a slot batch has idle slots
// without a network, so only the steady-state calls (which skip finished
// and idle columns) accept it.
int check_plain(vk_batch b, const char* who) {
  if (int rc = check_batch(b)) return rc;
  if (b->slots_mode) {
    vk_set_error("%s: slot batches are for the steady-state entry points only", who);
    return VK_ERR_ARG;
  }
  return VK_OK;
}

double ros2_gamma() { return 1.0 + 1.0 / sqrt(2.0); }

}  // namespace

extern "C" {

const char* vk_last_error(void) { return g_err; }
int vk_version(void) { return 100; }
int vk_max_species(void) { return VK_G * 7; }

// Longest-processing-time assignment of outputs to the VK_CHEM_NT lanes of a
// workgroup: costliest first, each to the least loaded lane (lowest lane on
// ties).  Each lane's outputs come back in ascending id order.
static std::vector<std::vector<int>> lpt_lanes(const std::vector<int>& cost) {
  std::vector<int> order(cost.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  using Load = std::pair<long, int>;
  std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
  for (int l = 0; l < VK_CHEM_NT; ++l) heap.push({0, l});
  std::vector<std::vector<int>> lanes(VK_CHEM_NT);
  for (int id : order) {
    Load x = heap.top();
    heap.pop();
    lanes[x.second].push_back(id);
    heap.push({x.first + cost[id], x.second});
  }
  for (auto& v : lanes) std::sort(v.begin(), v.end());
  return lanes;
}

// Lane-interleaved record table from per-lane record lists, padded with
// no-op records to the longest lane rounded up to whole load groups.  A
// record with no term points at zero_item, the table slot the kernel holds
// at 0.0, so every record adds c * (a table entry).  Returns the records per
// lane.
static int interleave(const std::vector<std::vector<VkRec>>& per_lane, std::vector<VkRec>& out,
                      uint32_t zero_item) {
  size_t k = 1;
  for (auto& v : per_lane) k = std::max(k, v.size());
  k = (k + VK_REC_GROUP - 1) / VK_REC_GROUP * VK_REC_GROUP;  // whole load groups
  out.assign(k * VK_CHEM_NT, VkRec{0.0f, vk_rec_word(zero_item, VK_REC_NONE_DEST)});
  for (int l = 0; l < VK_CHEM_NT; ++l)
    for (size_t j = 0; j < per_lane[l].size(); ++j) out[j * VK_CHEM_NT + l] = per_lane[l][j];
  return (int)k;
}

int vk_network_create(int ni, int nr, int T, const int64_t* reactant_idx,
                      const int64_t* product_idx, const double* reactant_stoich,
                      const double* product_stoich, const uint8_t* is_three_body,
                      vk_network* out) {
  g_err[0] = 0;
  if (!out || !reactant_idx || !product_idx || !reactant_stoich || !product_stoich) {
    vk_set_error("vk_network_create: null pointer");
    return VK_ERR_ARG;
  }
  if (ni < 1 || ni > vk_max_species() || nr < 0 || T < 1 || T > 4) {
    vk_set_error("vk_network_create: unsupported sizes ni=%d nr=%d T=%d (ni<=%d, T<=4)", ni, nr,
                 T, vk_max_species());
    return VK_ERR_UNSUPPORTED;
  }
  const int S = nr + 1;
  std::vector<int> r_idx(S * T), p_idx(S * T);
  std::vector<double> r_st(S * T), p_st(S * T);
  std::vector<int> is3(S, 0);
  for (int r = 0; r < S; ++r) {
    for (int s = 0; s < T; ++s) {
      const int64_t ri = reactant_idx[r * T + s], pi = product_idx[r * T + s];
      if (ri < 0 || ri > ni || pi < 0 || pi > ni) {
        vk_set_error("vk_network_create: species index out of range at row %d slot %d", r, s);
        return VK_ERR_ARG;
      }
      const double st = reactant_stoich[r * T + s];
      if (st < 0 || st != floor(st) || st > 8 || product_stoich[r * T + s] < 0) {
        vk_set_error("vk_network_create: stoichiometry must be a non-negative integer <= 8 "
                     "(row %d slot %d)", r, s);
        return VK_ERR_UNSUPPORTED;
      }
      r_idx[r * T + s] = (int)ri;
      r_st[r * T + s] = st;
      p_idx[r * T + s] = (int)pi;
      p_st[r * T + s] = product_stoich[r * T + s];
    }
    is3[r] = is_three_body ? (is_three_body[r] != 0) : 0;
  }
  // The chemistry record streams hold coefficients as float and pack row
  // indices in 18 bits, tile / chunk indices in 14.
  for (int e = 0; e < S * T; ++e) {
    if ((double)(float)r_st[e] != r_st[e] || (double)(float)p_st[e] != p_st[e]) {
      vk_set_error("vk_network_create: stoichiometric coefficient %g is not exact in float",
                   (double)(float)r_st[e] != r_st[e] ? r_st[e] : p_st[e]);
      return VK_ERR_UNSUPPORTED;
    }
  }
  if ((int64_t)S * T >= (int64_t)VK_REC_NONE_ITEM ||
      (int64_t)vk_slots(ni) * vk_slots(ni) * VK_NT >= (int64_t)VK_REC_NONE_DEST ||
      2 * (int64_t)S * T / VK_FCHUNK + 2 * ni >= (int64_t)VK_REC_NONE_DEST) {
    vk_set_error("vk_network_create: network too large for the record streams (S*T=%d)", S * T);
    return VK_ERR_UNSUPPORTED;
  }
  // species-major loss / production lists in ascending (row, slot) order
  std::vector<int> lp(ni + 1, 0), pp(ni + 1, 0);
  std::vector<std::vector<std::pair<int, double>>> L(ni), P(ni);
  for (int r = 0; r < S; ++r)
    for (int s = 0; s < T; ++s) {
      const int ri = (int)reactant_idx[r * T + s], pi = (int)product_idx[r * T + s];
      if (ri < ni) L[ri].push_back({r * T + s, reactant_stoich[r * T + s]});
      if (pi < ni) P[pi].push_back({r * T + s, product_stoich[r * T + s]});
    }
  std::vector<int> li, pi_;
  std::vector<double> lc, pc;
  for (int i = 0; i < ni; ++i) {
    lp[i] = (int)li.size();
    for (auto& e : L[i]) { li.push_back(e.first); lc.push_back(e.second); }
    pp[i] = (int)pi_.size();
    for (auto& e : P[i]) { pi_.push_back(e.first); pc.push_back(e.second); }
  }
  lp[ni] = (int)li.size();
  pp[ni] = (int)pi_.size();
  // entry-major Jacobian lists in ascending (row, out slot, in slot) order
  std::vector<std::vector<std::pair<int, double>>> JL((size_t)ni * ni);
  for (int r = 0; r < S; ++r)
    for (int so = 0; so < 2 * T; ++so) {
      const int oi = so < T ? (int)reactant_idx[r * T + so] : (int)product_idx[r * T + so - T];
      const double oc = so < T ? -reactant_stoich[r * T + so] : product_stoich[r * T + so - T];
      if (oi >= ni) continue;
      for (int sj = 0; sj < T; ++sj) {
        const int ii = (int)reactant_idx[r * T + sj];
        if (ii >= ni) continue;
        JL[(size_t)oi * ni + ii].push_back({r * T + sj, oc});
      }
    }
  std::vector<int> jp((size_t)ni * ni + 1), ji;
  std::vector<double> jc;
  for (size_t e = 0; e < JL.size(); ++e) {
    jp[e] = (int)ji.size();
    for (auto& x : JL[e]) { ji.push_back(x.first); jc.push_back(x.second); }
  }
  jp[JL.size()] = (int)ji.size();
  // Jacobian tile records: one output per entry of the SL x SL tile layout
  // (padding entries have no terms), each with its list in order.
  const int SL = vk_slots(ni);
  const int ntile = SL * SL * VK_NT;
  auto tile_elem = [&](int idx, int& i, int& q) {
    const int blk = idx / VK_NT, rc = idx % VK_NT;
    i = (blk / SL) * VK_G + rc / VK_G;
    q = (blk % SL) * VK_G + rc % VK_G;
  };
  std::vector<int> jcost(ntile);
  for (int idx = 0; idx < ntile; ++idx) {
    int i, q;
    tile_elem(idx, i, q);
    const size_t len = (i < ni && q < ni) ? JL[(size_t)i * ni + q].size() : 0;
    jcost[idx] = (int)std::max<size_t>(len, 1);
  }
  std::vector<std::vector<VkRec>> jl(VK_CHEM_NT);
  {
    const auto lanes = lpt_lanes(jcost);
    for (int l = 0; l < VK_CHEM_NT; ++l)
      for (int idx : lanes[l]) {
        int i, q;
        tile_elem(idx, i, q);
        if (i >= ni || q >= ni || JL[(size_t)i * ni + q].empty()) {
          jl[l].push_back({0.0f, vk_rec_word((uint32_t)(S * T), (uint32_t)idx)});
          continue;
        }
        const auto& lst = JL[(size_t)i * ni + q];
        for (size_t m = 0; m < lst.size(); ++m)
          jl[l].push_back({(float)lst[m].second,
                           vk_rec_word((uint32_t)lst[m].first,
                                       m + 1 == lst.size() ? (uint32_t)idx : VK_REC_NONE_DEST)});
      }
  }
  std::vector<VkRec> jrec;
  const int jrec_k = interleave(jl, jrec, (uint32_t)(S * T));
  // loss / production chunk records: loss chunks of species 0..ni-1, then
  // production chunks, VK_FCHUNK list entries each
  std::vector<int> lcp(ni + 1), pcp(ni + 1);
  std::vector<std::vector<VkRec>> chunks;
  auto add_chunks = [&](const std::vector<std::pair<int, double>>& lst) {
    for (size_t b = 0; b < lst.size(); b += VK_FCHUNK) {
      std::vector<VkRec> ch;
      const size_t e = std::min(lst.size(), b + VK_FCHUNK);
      for (size_t m = b; m < e; ++m)
        ch.push_back({(float)lst[m].second,
                      vk_rec_word((uint32_t)(lst[m].first / T),
                                  m + 1 == e ? (uint32_t)chunks.size() : VK_REC_NONE_DEST)});
      chunks.push_back(ch);
    }
  };
  for (int i = 0; i < ni; ++i) { lcp[i] = (int)chunks.size(); add_chunks(L[i]); }
  lcp[ni] = (int)chunks.size();
  for (int i = 0; i < ni; ++i) { pcp[i] = (int)chunks.size(); add_chunks(P[i]); }
  pcp[ni] = (int)chunks.size();
  const int nchunk = (int)chunks.size();
  std::vector<std::vector<VkRec>> fl(VK_CHEM_NT);
  {
    std::vector<int> fcost(nchunk);
    for (int ch = 0; ch < nchunk; ++ch) fcost[ch] = (int)chunks[ch].size();
    const auto lanes = lpt_lanes(fcost);
    for (int l = 0; l < VK_CHEM_NT; ++l)
      for (int ch : lanes[l]) fl[l].insert(fl[l].end(), chunks[ch].begin(), chunks[ch].end());
  }
  std::vector<VkRec> frec;
  const int frec_k = interleave(fl, frec, (uint32_t)S);
  // one device allocation for every table
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  const size_t o_pidx = take(p_idx.size() * 4), o_pst = take(p_st.size() * 8);
  const size_t o_ridx = take(r_idx.size() * 4), o_rst = take(r_st.size() * 8),
               o_is3 = take(is3.size() * 4), o_lp = take(lp.size() * 4),
               o_li = take(std::max<size_t>(li.size(), 1) * 4),
               o_lc = take(std::max<size_t>(lc.size(), 1) * 8), o_pp = take(pp.size() * 4),
               o_pi = take(std::max<size_t>(pi_.size(), 1) * 4),
               o_pc = take(std::max<size_t>(pc.size(), 1) * 8), o_jp = take(jp.size() * 4),
               o_ji = take(std::max<size_t>(ji.size(), 1) * 4),
               o_jc = take(std::max<size_t>(jc.size(), 1) * 8),
               o_jr = take(jrec.size() * sizeof(VkRec)), o_fr = take(frec.size() * sizeof(VkRec)),
               o_lcp = take(lcp.size() * 4), o_pcp = take(pcp.size() * 4);
  std::vector<char> host(off, 0);
  auto put = [&](size_t o, const void* src, size_t bytes) { if (bytes) memcpy(&host[o], src, bytes); };
  put(o_pidx, p_idx.data(), p_idx.size() * 4);
  put(o_pst, p_st.data(), p_st.size() * 8);
  put(o_ridx, r_idx.data(), r_idx.size() * 4);
  put(o_rst, r_st.data(), r_st.size() * 8);
  put(o_is3, is3.data(), is3.size() * 4);
  put(o_lp, lp.data(), lp.size() * 4);
  put(o_li, li.data(), li.size() * 4);
  put(o_lc, lc.data(), lc.size() * 8);
  put(o_pp, pp.data(), pp.size() * 4);
  put(o_pi, pi_.data(), pi_.size() * 4);
  put(o_pc, pc.data(), pc.size() * 8);
  put(o_jp, jp.data(), jp.size() * 4);
  put(o_ji, ji.data(), ji.size() * 4);
  put(o_jc, jc.data(), jc.size() * 8);
  put(o_jr, jrec.data(), jrec.size() * sizeof(VkRec));
  put(o_fr, frec.data(), frec.size() * sizeof(VkRec));
  put(o_lcp, lcp.data(), lcp.size() * 4);
  put(o_pcp, pcp.data(), pcp.size() * 4);
  vk_network_s* n = new vk_network_s();
  if (hipMalloc(&n->buf, off) != hipSuccess) {
    delete n;
    vk_set_error("vk_network_create: hipMalloc(%zu) failed", off);
    return VK_ERR_NOMEM;
  }
  if (hipMemcpy(n->buf, host.data(), off, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(n->buf);
    delete n;
    vk_set_error("vk_network_create: upload failed");
    return VK_ERR_HIP;
  }
  char* b = (char*)n->buf;
  n->dev.ni = ni; n->dev.S = S; n->dev.T = T; n->dev.pad = 0;
  n->dev.r_idx = (const int*)(b + o_ridx);
  n->dev.r_st = (const double*)(b + o_rst);
  n->dev.p_idx = (const int*)(b + o_pidx);
  n->dev.p_st = (const double*)(b + o_pst);
  n->dev.is3 = (const int*)(b + o_is3);
  n->dev.loss_ptr = (const int*)(b + o_lp);
  n->dev.loss_item = (const int*)(b + o_li);
  n->dev.loss_c = (const double*)(b + o_lc);
  n->dev.prod_ptr = (const int*)(b + o_pp);
  n->dev.prod_item = (const int*)(b + o_pi);
  n->dev.prod_c = (const double*)(b + o_pc);
  n->dev.jac_ptr = (const int*)(b + o_jp);
  n->dev.jac_item = (const int*)(b + o_ji);
  n->dev.jac_c = (const double*)(b + o_jc);
  n->dev.jrec = (const VkRec*)(b + o_jr);
  n->dev.frec = (const VkRec*)(b + o_fr);
  n->dev.lc_ptr = (const int*)(b + o_lcp);
  n->dev.pc_ptr = (const int*)(b + o_pcp);
  n->dev.jrec_k = jrec_k;
  n->dev.frec_k = frec_k;
  n->dev.nchunk = nchunk;
  n->dev.pad2 = 0;
  n->ni = ni; n->nr = nr; n->T = T;
  *out = n;
  return VK_OK;
}

int vk_network_destroy(vk_network net) {
  if (!net) return VK_OK;
  if (net->buf) (void)hipFree(net->buf);
  delete net;
  return VK_OK;
}

int64_t vk_block_doubles(int ni) {
  const int sl = vk_slots(ni);
  return (int64_t)sl * sl * VK_NT;
}

int vk_batch_create(int ncols, const int32_t* nz, const int32_t* ni, const int32_t* S,
                    const int32_t* net, const int64_t* off_s, const int64_t* off_k,
                    const int64_t* off_a, const int64_t* off_b, const vk_network* nets,
                    int nnets, vk_batch* out) {
  g_err[0] = 0;
  if (!out || ncols < 1 || !nz || !ni || !S || !net || !off_s || !off_k || !off_a || !off_b ||
      !nets || nnets < 1) {
    vk_set_error("vk_batch_create: bad arguments");
    return VK_ERR_ARG;
  }
  vk_batch_s* b = new vk_batch_s();
  b->ncols = ncols;
  b->cols.resize(ncols);
  std::set<int> sls;
  for (int c = 0; c < ncols; ++c) {
    if (net[c] < 0 || net[c] >= nnets || !nets[net[c]]) {
      delete b;
      vk_set_error("vk_batch_create: column %d names network %d of %d", c, net[c], nnets);
      return VK_ERR_ARG;
    }
    const vk_network_s* nw = nets[net[c]];
    if (nz[c] < 2 || ni[c] != nw->ni || S[c] != nw->nr + 1 || ni[c] > vk_max_species()) {
      delete b;
      vk_set_error("vk_batch_create: column %d has nz=%d ni=%d S=%d but its network has "
                   "ni=%d S=%d (nz >= 2, ni <= %d)", c, nz[c], ni[c], S[c], nw->ni, nw->nr + 1,
                   vk_max_species());
      return VK_ERR_ARG;
    }
    VkCol& cl = b->cols[c];
    cl.nz = nz[c]; cl.ni = ni[c]; cl.S = S[c]; cl.net = net[c];
    cl.off_s = off_s[c]; cl.off_k = off_k[c]; cl.off_a = off_a[c]; cl.off_b = off_b[c];
    sls.insert(vk_slots(ni[c]));
    b->max_nz = std::max(b->max_nz, (int)nz[c]);
    b->max_ni = std::max(b->max_ni, (int)ni[c]);
    b->max_rec = std::max(b->max_rec, vk_atm_offsets((int)nz[c], (int)ni[c]).total);
    b->max_S = std::max(b->max_S, (int)S[c]);
    b->max_T = std::max(b->max_T, nw->T);
    b->state_doubles = std::max<int64_t>(b->state_doubles, off_s[c] + (int64_t)nz[c] * ni[c]);
    b->k_doubles = std::max<int64_t>(b->k_doubles, off_k[c] + (int64_t)S[c] * nz[c]);
    b->block_doubles =
        std::max<int64_t>(b->block_doubles, off_b[c] + (int64_t)nz[c] * vk_block_doubles(ni[c]));
  }
  b->slots.assign(sls.begin(), sls.end());
  std::vector<VkNetDev> hn(nnets);
  for (int i = 0; i < nnets; ++i) hn[i] = nets[i] ? nets[i]->dev : VkNetDev{};
  b->net_dims.resize(nnets, {-1, -1});
  for (int i = 0; i < nnets; ++i)
    if (nets[i]) b->net_dims[i] = {nets[i]->ni, nets[i]->nr};
  if (hipMalloc(&b->d_cols, sizeof(VkCol) * ncols) != hipSuccess ||
      hipMalloc(&b->d_nets, sizeof(VkNetDev) * nnets) != hipSuccess) {
    vk_batch_destroy(b);
    vk_set_error("vk_batch_create: hipMalloc failed");
    return VK_ERR_NOMEM;
  }
  if (hipMemcpy(b->d_cols, b->cols.data(), sizeof(VkCol) * ncols, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_nets, hn.data(), sizeof(VkNetDev) * nnets, hipMemcpyHostToDevice) !=
          hipSuccess) {
    vk_batch_destroy(b);
    vk_set_error("vk_batch_create: upload failed");
    return VK_ERR_HIP;
  }
  *out = b;
  return VK_OK;
}

int vk_batch_uniform(int ncols, int nz, vk_network net, vk_batch* out) {
  if (!net) { vk_set_error("vk_batch_uniform: null network"); return VK_ERR_ARG; }
  std::vector<int32_t> vnz(ncols, nz), vni(ncols, net->ni), vS(ncols, net->nr + 1),
      vnet(ncols, 0);
  std::vector<int64_t> os(ncols), ok(ncols), oa(ncols), ob(ncols);
  const int64_t alen = vk_atm_offsets(nz, net->ni).total;
  for (int c = 0; c < ncols; ++c) {
    os[c] = (int64_t)c * nz * net->ni;
    ok[c] = (int64_t)c * (net->nr + 1) * nz;
    oa[c] = (int64_t)c * alen;
    ob[c] = (int64_t)c * nz * vk_block_doubles(net->ni);
  }
  vk_network nets[1] = {net};
  int rc = vk_batch_create(ncols, vnz.data(), vni.data(), vS.data(), vnet.data(), os.data(),
                           ok.data(), oa.data(), ob.data(), nets, 1, out);
  if (rc == VK_OK) (*out)->uniform = true;
  return rc;
}

__global__ void k_set_slot(VkCol* cols, VkNetDev* nets, int slot, VkCol col, VkNetDev net) {
  cols[slot] = col;
  nets[slot] = net;
}

int vk_batch_slots(int nslots, int cap_nz, int cap_ni, int cap_S, int cap_T, vk_batch* out) {
  g_err[0] = 0;
  if (!out || nslots < 1 || cap_nz < 2 || cap_ni < 1 || cap_ni > vk_max_species() || cap_S < 1 ||
      cap_T < 1) {
    vk_set_error("vk_batch_slots: bad arguments (nslots=%d cap_nz=%d cap_ni=%d cap_S=%d cap_T=%d)",
                 nslots, cap_nz, cap_ni, cap_S, cap_T);
    return VK_ERR_ARG;
  }
  vk_batch_s* b = new vk_batch_s();
  b->ncols = nslots;
  b->slots_mode = true;
  b->cap_nz = cap_nz; b->cap_ni = cap_ni; b->cap_S = cap_S; b->cap_T = cap_T;
  b->cols.resize(nslots);
  const int64_t cs = (int64_t)cap_nz * cap_ni, ck = (int64_t)cap_S * cap_nz;
  const int64_t ca = vk_atm_offsets(cap_nz, cap_ni).total;
  const int64_t cb = (int64_t)cap_nz * vk_block_doubles(cap_ni);
  for (int c = 0; c < nslots; ++c) {
    VkCol& cl = b->cols[c];
    cl.nz = cap_nz; cl.ni = cap_ni; cl.S = cap_S; cl.net = c;
    cl.off_s = c * cs; cl.off_k = c * ck; cl.off_a = c * ca; cl.off_b = c * cb;
  }
  b->slots.assign(1, vk_slots(cap_ni));
  b->max_nz = cap_nz; b->max_ni = cap_ni; b->max_S = cap_S; b->max_T = cap_T;
  b->max_rec = ca;
  b->state_doubles = nslots * cs;
  b->k_doubles = nslots * ck;
  b->block_doubles = nslots * cb;
  std::vector<VkNetDev> hn(nslots, VkNetDev{});
  if (hipMalloc(&b->d_cols, sizeof(VkCol) * nslots) != hipSuccess ||
      hipMalloc(&b->d_nets, sizeof(VkNetDev) * nslots) != hipSuccess) {
    vk_batch_destroy(b);
    vk_set_error("vk_batch_slots: hipMalloc failed");
    return VK_ERR_NOMEM;
  }
  if (hipMemcpy(b->d_cols, b->cols.data(), sizeof(VkCol) * nslots, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_nets, hn.data(), sizeof(VkNetDev) * nslots, hipMemcpyHostToDevice) !=
          hipSuccess) {
    vk_batch_destroy(b);
    vk_set_error("vk_batch_slots: upload failed");
    return VK_ERR_HIP;
  }
  *out = b;
  return VK_OK;
}

int vk_batch_set_column(vk_batch b, int slot, int nz, vk_network net, void* stream) {
  g_err[0] = 0;
  if (int rc = check_batch(b)) return rc;
  if (!b->slots_mode) { vk_set_error("vk_batch_set_column: not a slot batch"); return VK_ERR_ARG; }
  if (slot < 0 || slot >= b->ncols || !net) {
    vk_set_error("vk_batch_set_column: slot %d of %d or null network", slot, b->ncols);
    return VK_ERR_ARG;
  }
  if (nz < 2 || nz > b->cap_nz || net->ni > b->cap_ni || vk_slots(net->ni) != vk_slots(b->cap_ni) ||
      net->nr + 1 > b->cap_S || net->T > b->cap_T) {
    vk_set_error("vk_batch_set_column: nz=%d ni=%d S=%d T=%d does not fit the slot (nz<=%d, "
                 "ni<=%d in the same 16-species bucket, S<=%d, T<=%d)", nz, net->ni, net->nr + 1,
                 net->T, b->cap_nz, b->cap_ni, b->cap_S, b->cap_T);
    return VK_ERR_ARG;
  }
  VkCol cl = b->cols[slot];
  cl.nz = nz; cl.ni = net->ni; cl.S = net->nr + 1; cl.net = slot;
  b->cols[slot] = cl;
  hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(1), 0, (hipStream_t)stream, b->d_cols, b->d_nets,
                     slot, cl, net->dev);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_batch_destroy(vk_batch b) {
  if (!b) return VK_OK;
  for (hipStream_t a : b->aux) (void)hipStreamDestroy(a);
  for (hipEvent_t e : b->ev_join) (void)hipEventDestroy(e);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->d_cols) (void)hipFree(b->d_cols);
  if (b->d_nets) (void)hipFree(b->d_nets);
  if (b->d_rates) (void)hipFree(b->d_rates);
  delete b;
  return VK_OK;
}

int64_t vk_batch_block_doubles(vk_batch b) { return b ? b->block_doubles : 0; }

int64_t vk_atm_record_layout(int nz, int ni, int64_t* off) {
  const VkAtmOff o = vk_atm_offsets(nz, ni);
  if (off) {
    const int64_t v[18] = {o.Kzz, o.dzi, o.vz, o.Hpi, o.Ti, o.Tco, o.g, o.M, o.ms,
                           o.alpha, o.top_flux, o.bot_flux, o.bot_vdep, o.gas, o.Dzz, o.vm,
                           o.vs, o.flags};
    memcpy(off, v, sizeof(v));
  }
  return o.total;
}

static int chem_call(vk_batch b, int mode, const double* y, const double* M, const double* atm,
                     const double* k, const double* dt, const uint8_t* fix, const double* k1,
                     double* out0, double* out_blk, double* sup, double* sub, const int* active,
                     hipStream_t st) {
  for (int sl : b->slots) {
    int rc = vk_launch_chem_layer(sl, mode, b->d_cols, b->ncols, b->max_nz, b->max_S, b->max_T,
                                  b->d_nets, y, M, atm, k, dt, fix, k1, out0, out_blk, sup, sub,
                                  active, 0, nullptr, 0, st);
    if (rc) return rc;
  }
  return VK_OK;
}

int vk_chem_rhs_f64(vk_batch b, const double* y, const double* M, const double* atm,
                    const double* k, double* out, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_chem_rhs_f64")) return rc;
  if (!y || !k || !out || (!M && !atm)) { vk_set_error("vk_chem_rhs_f64: null pointer"); return VK_ERR_ARG; }
  if (M && !b->uniform) { vk_set_error("vk_chem_rhs_f64: explicit M needs a uniform batch"); return VK_ERR_ARG; }
  return chem_call(b, M_CHEM_RHS, y, M, atm, k, nullptr, nullptr, nullptr, out, nullptr, nullptr,
                   nullptr, nullptr, (hipStream_t)stream);
}

int vk_chem_jac_f64(vk_batch b, const double* y, const double* M, const double* atm,
                    const double* k, double* jac, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_chem_jac_f64")) return rc;
  if (!y || !k || !jac || (!M && !atm)) { vk_set_error("vk_chem_jac_f64: null pointer"); return VK_ERR_ARG; }
  if (M && !b->uniform) { vk_set_error("vk_chem_jac_f64: explicit M needs a uniform batch"); return VK_ERR_ARG; }
  return chem_call(b, M_CHEM_JAC, y, M, atm, k, nullptr, nullptr, nullptr, jac, nullptr, nullptr,
                   nullptr, nullptr, (hipStream_t)stream);
}

int vk_residual_f64(vk_batch b, const double* y, const double* k, const double* atm, double* out,
                    void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_residual_f64")) return rc;
  if (!y || !k || !atm || !out) { vk_set_error("vk_residual_f64: null pointer"); return VK_ERR_ARG; }
  return chem_call(b, M_FTOT, y, nullptr, atm, k, nullptr, nullptr, nullptr, out, nullptr, nullptr,
                   nullptr, nullptr, (hipStream_t)stream);
}

int vk_jacobian_blocks_f64(vk_batch b, const double* y, const double* k, const double* atm,
                           double* diag, double* sup_d, double* sub_d, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_jacobian_blocks_f64")) return rc;
  if (!y || !k || !atm || !diag || !sup_d || !sub_d) {
    vk_set_error("vk_jacobian_blocks_f64: null pointer");
    return VK_ERR_ARG;
  }
  return chem_call(b, M_JY, y, nullptr, atm, k, nullptr, nullptr, nullptr, diag, nullptr, sup_d,
                   sub_d, nullptr, (hipStream_t)stream);
}

int vk_bt_factor_f64(vk_batch b, const double* diag, const double* sup_d, const double* sub_d,
                     int transpose, double* inv, int32_t* status, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_bt_factor_f64")) return rc;
  if (!diag || !sup_d || !sub_d || !inv) { vk_set_error("vk_bt_factor_f64: null pointer"); return VK_ERR_ARG; }
  for (int sl : b->slots) {
    int rc = vk_launch_bt_factor(sl, b->d_cols, b->ncols, diag, 0, transpose, sup_d, sub_d, inv,
                                 nullptr, nullptr, (int*)status, nullptr, (hipStream_t)stream);
    if (rc) return rc;
  }
  return VK_OK;
}

size_t vk_bt_cr_workspace_bytes(vk_batch b) {
  return b ? (size_t)(8 * b->block_doubles) * sizeof(double) : 0;
}

int vk_bt_cr_factor_f64(vk_batch b, const double* diag, const double* sup_d, const double* sub_d,
                        int transpose, void* ws, size_t ws_bytes, int32_t* status, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_bt_cr_factor_f64")) return rc;
  if (!diag || !sup_d || !sub_d || !ws || ws_bytes < vk_bt_cr_workspace_bytes(b)) {
    vk_set_error("vk_bt_cr_factor_f64: null pointer or workspace %zu < %zu bytes", ws_bytes,
                 vk_bt_cr_workspace_bytes(b));
    return VK_ERR_ARG;
  }
  double* tile = (double*)ws;
  double* cr = tile + b->block_doubles;
  hipStream_t st = (hipStream_t)stream;
  for (int sl : b->slots) {
    if (int rc = vk_launch_nat_to_tile(sl, b->d_cols, b->ncols, b->max_nz, diag, tile,
                                       transpose, st))
      return rc;
    if (int rc = vk_launch_cr_factor(sl, b->d_cols, b->ncols, b->max_nz, tile, sup_d, sub_d, cr,
                                     (int*)status, nullptr, st))
      return rc;
  }
  return VK_OK;
}

int vk_bt_cr_solve_f64(vk_batch b, const double* sup_d, const double* sub_d, const double* rhs,
                       double* x, void* ws, size_t ws_bytes, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_bt_cr_solve_f64")) return rc;
  if (!sup_d || !sub_d || !rhs || !x || !ws || ws_bytes < vk_bt_cr_workspace_bytes(b)) {
    vk_set_error("vk_bt_cr_solve_f64: null pointer or workspace %zu < %zu bytes", ws_bytes,
                 vk_bt_cr_workspace_bytes(b));
    return VK_ERR_ARG;
  }
  double* cr = (double*)ws + b->block_doubles;
  SolveEpi epi{};
  for (int sl : b->slots)
    if (int rc = vk_launch_cr_solve(sl, b->d_cols, b->ncols, b->max_nz, rhs, sup_d, sub_d, cr, x,
                                    epi, nullptr, (hipStream_t)stream))
      return rc;
  return VK_OK;
}

int vk_bt_solve_f64(vk_batch b, const double* inv, const double* sup_d, const double* sub_d,
                    const double* rhs, double* x, double* scratch, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_bt_solve_f64")) return rc;
  if (!inv || !sup_d || !sub_d || !rhs || !x || !scratch) {
    vk_set_error("vk_bt_solve_f64: null pointer");
    return VK_ERR_ARG;
  }
  SolveEpi epi{};
  for (int sl : b->slots) {
    int rc = vk_launch_bt_solve(sl, b->d_cols, b->ncols, inv, sup_d, sub_d, rhs, scratch, x, 0,
                                epi, nullptr, (hipStream_t)stream);
    if (rc) return rc;
  }
  return VK_OK;
}

int vk_diff_grav_f64(vk_batch b, const double* atm, double* A, double* B, double* C,
                     void* stream) {
  g_err[0] = 0;
  if (int rc = check_batch(b)) return rc;
  if (!atm || !A || !B || !C) {
    vk_set_error("vk_diff_grav_f64: null pointer");
    return VK_ERR_ARG;
  }
  return vk_launch_diff_grav(b->d_cols, b->ncols, b->max_nz, atm, A, B, C, (hipStream_t)stream);
}

int vk_bt_dense_f64(vk_batch b, const double* diag, const double* sup, const double* sub,
                    const double* rhs, double* x, double* inv, double* scratch, int32_t* status,
                    void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_bt_dense_f64")) return rc;
  if (!diag || !sup || !sub || !rhs || !x || !inv || !scratch) {
    vk_set_error("vk_bt_dense_f64: null pointer");
    return VK_ERR_ARG;
  }
  for (int sl : b->slots) {
    int rc = vk_launch_bt_dense(sl, b->d_cols, b->ncols, diag, sup, sub, rhs, x, inv, scratch,
                                (int*)status, (hipStream_t)stream);
    if (rc) return rc;
  }
  return VK_OK;
}

int vk_bt_inverse_natural_f64(vk_batch b, const double* inv, double* nat, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_bt_inverse_natural_f64")) return rc;
  for (int sl : b->slots) {
    int rc = vk_launch_tile_to_nat(sl, b->d_cols, b->ncols, b->max_nz, inv, nat,
                                   (hipStream_t)stream);
    if (rc) return rc;
  }
  return VK_OK;
}

int vk_timing_enable(int on) {
  g_timing.on = on != 0;
  return VK_OK;
}

int vk_timing_reset(void) {
  for (auto& p : g_timing.pending) {
    g_timing.pool.push_back(p.a);
    g_timing.pool.push_back(p.b);
  }
  g_timing.pending.clear();
  for (int i = 0; i < VK_NKERN; ++i) { g_timing.ms[i] = 0; g_timing.count[i] = 0; }
  return VK_OK;
}

int vk_timing_read(double* ms, long long* count, int n) {
  for (auto& p : g_timing.pending) {
    float t = 0.f;
    VK_CHECK_HIP(hipEventSynchronize(p.b));
    VK_CHECK_HIP(hipEventElapsedTime(&t, p.a, p.b));
    g_timing.ms[p.kind] += t;
    g_timing.count[p.kind] += 1;
    g_timing.pool.push_back(p.a);
    g_timing.pool.push_back(p.b);
  }
  g_timing.pending.clear();
  for (int i = 0; i < n && i < VK_NKERN; ++i) {
    if (ms) ms[i] = g_timing.ms[i];
    if (count) count[i] = g_timing.count[i];
  }
  return VK_OK;
}

// VK_TUNE_MOLPRE: 1 (default) computes the y-independent molecular-diffusion
// factors once per step (k_mol_static); 0 forms them inside each chemistry
// launch.  The coefficients are the same bit for bit either way.
static int g_mol_precompute = 1;

// VK_TUNE_REFINE: 0 none; 1 the stage-1 solve is refined; 2 (default) both
// stages' solves.  The block-Thomas factors refine each block solve of the
// backward sweep, and a column where some block's correction exceeded the
// solution (an explicit inverse with no correct digit) takes one
// whole-column step on top; the cyclic-reduction factor takes one
// whole-column refinement step (residual, correction solve) for every
// column.
static int g_refine = 2;

// Offset of the inverse stack in the step workspace: the factor writes the
// inverses there, so the diagonal blocks D survive for the refinement.
static size_t ros2_ws_inv_offset(vk_batch b) {
  return align_up<double>(b->block_doubles) + 16 * align_up<double>(b->state_doubles) +
         align_up<double>(b->k_doubles);
}

// Offset of the Schur-complement stack: with refinement the tile factor
// stores each A'_j there, and the refined solves take their residuals on it.
static size_t ros2_ws_schur_offset(vk_batch b) {
  return ros2_ws_inv_offset(b) + align_up<double>(b->block_doubles);
}

// Offset of the per-column whole-column refinement flags (ints).
static size_t ros2_ws_refl_offset(vk_batch b) {
  return ros2_ws_schur_offset(b) + align_up<double>(b->block_doubles);
}

// Bytes before the cyclic-reduction region of the step workspace.
static size_t ros2_ws_base_bytes(vk_batch b) {
  return ros2_ws_refl_offset(b) + align_up<int>((size_t)b->ncols);
}

// Whether bucket sl of b uses the cyclic-reduction factor.
static bool bucket_is_cr(vk_batch b, int sl) { return b->latency_mode || vk_factor_is_cr(sl); }

// Whether a bucket of b uses the cyclic-reduction factor (it then needs 7
// block stacks of extra workspace).
static bool batch_uses_cr(vk_batch b) {
  for (int sl : b->slots)
    if (bucket_is_cr(b, sl)) return true;
  return false;
}

int vk_batch_set_latency_mode(vk_batch b, int on) {
  g_err[0] = 0;
  if (int rc = check_batch(b)) return rc;
  if (on != 0 && on != 1) {
    vk_set_error("vk_batch_set_latency_mode: on must be 0 or 1, not %d", on);
    return VK_ERR_ARG;
  }
  b->latency_mode = on != 0;
  return VK_OK;
}

size_t vk_ros2_workspace_bytes(vk_batch b) {
  if (!b) return 0;
  return ros2_ws_base_bytes(b) + (batch_uses_cr(b) ? align_up<double>(7 * b->block_doubles) : 0);
}

// Helper streams and events for running the SL buckets of a batch
// concurrently; created on first use.  false: run the buckets in order.
static bool ensure_aux(vk_batch b, size_t n) {
  if (n < 2) return false;
  if (b->aux.size() >= n) return true;
  for (size_t i = b->aux.size(); i < n; ++i) {
    hipStream_t a = nullptr;
    hipEvent_t e = nullptr;
    if (hipStreamCreateWithFlags(&a, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      if (a) (void)hipStreamDestroy(a);
      return false;
    }
    b->aux.push_back(a);
    b->ev_join.push_back(e);
  }
  if (!b->ev_fork && hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

// Column chunks per size bucket in the Ros2 step (VK_TUNE_SPLIT): each
// chunk runs the bucket's kernel sequence on its own stream, so the
// latency-bound factor of one chunk overlaps the bandwidth-bound chemistry
// and sweeps of another.  A chunk holds at least kMinChunk columns.
static int g_split = 1;
constexpr int kMinChunk = 128;

// VK_TUNE_KDIRECT: 1 has the chemistry read the rate tables in their (S, nz)
// layout; 0 (default) copies them layer-major once per step first.
static int g_k_direct = 0;

static int ros2_step_impl(vk_batch b, const double* y, const double* k, const double* atm,
                          const double* dt, const uint8_t* fix_mask, double* sol, double* delta,
                          void* ws, int32_t* status, const int* active, hipStream_t st,
                          bool fork = true) {
  char* w = (char*)ws;
  double* blk = (double*)w; w += align_up<double>(b->block_doubles);
  const size_t sv = align_up<double>(b->state_doubles);
  double* rhs_y = (double*)w; w += sv;
  double* rmod = (double*)w; w += sv;
  double* k1 = (double*)w; w += sv;
  double* yk2 = (double*)w; w += sv;
  double* rhs2 = (double*)w; w += sv;
  double* k2 = (double*)w; w += sv;
  double* sup = (double*)w; w += sv;
  double* sub = (double*)w; w += sv;
  double* kT = (double*)w;  // the rate tables layer-major, read contiguously by the chemistry
  w += align_up<double>(b->k_doubles);
  double* mst = (double*)w;  // 8 components of the molecular-diffusion static factors
  const int64_t mstride = (int64_t)(sv / sizeof(double));
  double* crw = (double*)((char*)ws + ros2_ws_base_bytes(b));  // cyclic-reduction levels
  // with refinement the factor writes the inverses X_j to the second block
  // stack and the diagonal blocks D_j stay for the refined solves; without,
  // the inverses overwrite D_j
  double* inv = g_refine > 0 ? (double*)((char*)ws + ros2_ws_inv_offset(b)) : blk;
  const double* ref1 = g_refine >= 1 ? blk : nullptr;
  const double* ref2 = g_refine >= 2 ? blk : nullptr;
  double* schur = (double*)((char*)ws + ros2_ws_schur_offset(b));
  int* refl = (int*)((char*)ws + ros2_ws_refl_offset(b));
  const double* kchem = g_k_direct ? k : kT;  // the rate tables the chemistry reads
  const int klm = g_k_direct ? 0 : 1;
  const double r = ros2_gamma();
  {
    KernelTimer t(6, st);
    if (!g_k_direct) {
      if (int rc = vk_launch_rates_layer_major(b->d_cols, b->ncols, b->max_S, b->max_nz, k, kT,
                                               active, st))
        return rc;
    }
    if (g_mol_precompute) {
      if (int rc = vk_launch_mol_static(b->d_cols, b->ncols, b->max_nz, atm, mst, mstride,
                                        active, st))
        return rc;
    } else {
      mst = nullptr;
    }
  }
  // The kernel sequence of one SL bucket over the columns [c0, c0 + nc) on
  // stream s.  Every per-column array the kernels index by position (the
  // descriptors, dt, status, active) is offset to the range; state arrays
  // are addressed through the descriptors' offsets.
  auto run_bucket = [&](int sl, int c0, int nc, hipStream_t s) -> int {
    int rc;
    const VkCol* cols = b->d_cols + c0;
    const double* dtc = dt + c0;
    int32_t* stc = status ? status + c0 : nullptr;
    const int* act = active ? active + c0 : nullptr;
    int* rfl = refl + c0;
    {
      KernelTimer t(0, s);
      if ((rc = vk_launch_chem_layer(sl, M_LHS, cols, nc, b->max_nz, b->max_S,
                                     b->max_T, b->d_nets, y, nullptr, atm, kchem, dtc, fix_mask,
                                     nullptr, rhs_y, blk, sup, sub, act, klm, mst, mstride, s)))
        return rc;
    }
    // The factor fuses the stage-1 forward sweep unless the blocked variant
    // is selected; then the stage-1 solve runs both sweeps.  The
    // cyclic-reduction variant has its own factor and solves.
    const bool cr = bucket_is_cr(b, sl);
    const bool fused = !cr && vk_factor_fuses_sweep(sl);
    // refined block solves on the stored Schur complements (the tile factor)
    // or on D_j and X_{j-1} (the other variants)
    const bool sch = !cr && ref1 && vk_factor_stores_schur(sl);
    const double* rd1 = sch ? nullptr : ref1;
    const double* rd2 = sch ? nullptr : ref2;
    const double* rs1 = sch ? schur : nullptr;
    const double* rs2 = sch && ref2 ? schur : nullptr;
    {
      KernelTimer t(1, s);
      if (cr)  // the stage-1 reduction of rhs_y rides in the update launches
        rc = vk_launch_cr_factor(sl, cols, nc, b->max_nz, blk, sup, sub, crw,
                                 (int*)stc, act, s, rhs_y);
      else
        rc = vk_launch_bt_factor(sl, cols, nc, blk, 1, 0, sup, sub, inv,
                                 fused ? rhs_y : nullptr, fused ? rmod : nullptr, (int*)stc,
                                 act, s, sch ? schur : nullptr);
      if (rc) return rc;
    }
    SolveEpi e1{1, r, 0.0, y, nullptr, nullptr, yk2, nullptr};
    {
      KernelTimer t(2, s);
      if (cr && ref1) {
        // x0 into k2 (free until stage 2), its residual into rmod, then the
        // correction solve writes k1 = x0 + dx and yk2
        SolveEpi e0{};
        if ((rc = vk_launch_cr_solve(sl, cols, nc, b->max_nz, rhs_y, sup, sub, crw,
                                     k2, e0, act, s, true)))
          return rc;
        if ((rc = vk_launch_bt_residual(sl, cols, nc, b->max_nz, blk, sup, sub, rhs_y,
                                        k2, rmod, act, s)))
          return rc;
        SolveEpi e1r = e1;
        e1r.base = k2;
        rc = vk_launch_cr_solve(sl, cols, nc, b->max_nz, rmod, sup, sub, crw, k1, e1r,
                                act, s);
      } else if (cr) {
        rc = vk_launch_cr_solve(sl, cols, nc, b->max_nz, rhs_y, sup, sub, crw, k1,
                                e1, act, s, true);
      } else if (ref1) {
        // the columns rfl flags: residual into rhs2 (free until the stage-2
        // chemistry), then k1 += the correction, yk2 again
        if ((rc = vk_launch_bt_solve(sl, cols, nc, inv, sup, sub, fused ? rmod : rhs_y,
                                     rmod, k1, fused ? 1 : 0, e1, act, s, rd1, rfl, rs1)))
          return rc;
        if ((rc = vk_launch_bt_residual(sl, cols, nc, b->max_nz, blk, sup, sub, rhs_y,
                                        k1, rhs2, rfl, s)))
          return rc;
        SolveEpi e1r = e1;
        e1r.base = k1;
        rc = vk_launch_bt_solve(sl, cols, nc, inv, sup, sub, rhs2, rmod, k1, 0, e1r, rfl, s,
                                rd1, nullptr, rs1);
      } else {
        rc = vk_launch_bt_solve(sl, cols, nc, inv, sup, sub, fused ? rmod : rhs_y,
                                rmod, k1, fused ? 1 : 0, e1, act, s);
      }
      if (rc) return rc;
    }
    {
      KernelTimer t(3, s);
      if ((rc = vk_launch_chem_layer(sl, M_RHS2, cols, nc, b->max_nz, b->max_S,
                                     b->max_T, b->d_nets, yk2, nullptr, atm, kchem, dtc, fix_mask,
                                     k1, rhs2, nullptr, nullptr, nullptr, act, klm, mst, mstride,
                                     s)))
        return rc;
    }
    SolveEpi e2{2, 3.0 / (2.0 * r), 1.0 / (2.0 * r), y, k1, yk2, sol, delta};
    {
      KernelTimer t(4, s);
      if (cr && ref2) {
        // x0 into rmod, its residual into rhs_y (free after stage 1), then
        // the correction solve writes k2 = x0 + dx, sol and delta
        SolveEpi e0{};
        if ((rc = vk_launch_cr_solve(sl, cols, nc, b->max_nz, rhs2, sup, sub, crw,
                                     rmod, e0, act, s)))
          return rc;
        if ((rc = vk_launch_bt_residual(sl, cols, nc, b->max_nz, blk, sup, sub, rhs2,
                                        rmod, rhs_y, act, s)))
          return rc;
        SolveEpi e2r = e2;
        e2r.base = rmod;
        rc = vk_launch_cr_solve(sl, cols, nc, b->max_nz, rhs_y, sup, sub, crw, k2, e2r,
                                act, s);
      } else if (cr) {
        rc = vk_launch_cr_solve(sl, cols, nc, b->max_nz, rhs2, sup, sub, crw, k2, e2,
                                act, s);
      } else if (ref2) {
        // the columns rfl flags: residual into rhs_y (free after stage 1),
        // then k2 += the correction, sol and delta again
        if ((rc = vk_launch_bt_solve(sl, cols, nc, inv, sup, sub, rhs2, rmod, k2, 0, e2,
                                     act, s, rd2, rfl, rs2)))
          return rc;
        if ((rc = vk_launch_bt_residual(sl, cols, nc, b->max_nz, blk, sup, sub, rhs2,
                                        k2, rhs_y, rfl, s)))
          return rc;
        SolveEpi e2r = e2;
        e2r.base = k2;
        rc = vk_launch_bt_solve(sl, cols, nc, inv, sup, sub, rhs_y, rmod, k2, 0, e2r, rfl, s,
                                rd2, nullptr, rs2);
      } else {
        rc = vk_launch_bt_solve(sl, cols, nc, inv, sup, sub, rhs2, rmod, k2, 0, e2, act, s);
      }
      if (rc) return rc;
    }
    return VK_OK;
  };
  // column chunks: contiguous index ranges, each running every bucket (a
  // kernel returns at once for a column of another bucket)
  const int nchunk = std::max(1, std::min(g_split, b->ncols / kMinChunk));
  const size_t nstreams = (size_t)nchunk * b->slots.size();
  auto chunk = [&](int i, int* c0, int* nc) {
    *c0 = (int)((int64_t)b->ncols * i / nchunk);
    *nc = (int)((int64_t)b->ncols * (i + 1) / nchunk) - *c0;
  };
  if (fork && ensure_aux(b, nstreams)) {
    // fork: every stream waits for the work before this point on st
    VK_CHECK_HIP(hipEventRecord(b->ev_fork, st));
    size_t i = 0;
    for (int ch = 0; ch < nchunk; ++ch) {
      int c0, nc;
      chunk(ch, &c0, &nc);
      for (int sl : b->slots) {
        VK_CHECK_HIP(hipStreamWaitEvent(b->aux[i], b->ev_fork, 0));
        if (int rc = run_bucket(sl, c0, nc, b->aux[i])) return rc;
        VK_CHECK_HIP(hipEventRecord(b->ev_join[i], b->aux[i]));
        ++i;
      }
    }
    // join: st waits for every stream
    for (size_t j = 0; j < nstreams; ++j) VK_CHECK_HIP(hipStreamWaitEvent(st, b->ev_join[j], 0));
    return VK_OK;
  }
  for (int sl : b->slots)
    if (int rc = run_bucket(sl, 0, b->ncols, st)) return rc;
  return VK_OK;
}

int vk_ros2_step_f64(vk_batch b, const double* y, const double* k, const double* atm,
                     const double* dt, const uint8_t* fix_mask, double* sol, double* delta,
                     void* ws, size_t ws_bytes, int32_t* status, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_ros2_step_f64")) return rc;
  if (!y || !k || !atm || !dt || !sol || !delta || !ws) {
    vk_set_error("vk_ros2_step_f64: null pointer");
    return VK_ERR_ARG;
  }
  if (ws_bytes < vk_ros2_workspace_bytes(b)) {
    vk_set_error("vk_ros2_step_f64: workspace %zu < %zu bytes", ws_bytes, vk_ros2_workspace_bytes(b));
    return VK_ERR_ARG;
  }
  return ros2_step_impl(b, y, k, atm, dt, fix_mask, sol, delta, ws, status, nullptr,
                        (hipStream_t)stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The allocation-free Ros2 step.  Its workspace is a file-scope device
// array; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (the array's name is a host-side stand-in,
// not a device address).  Each device holds its own copy until the process
// ends; every call on a device shares it, so calls that overlap on one
// device race.  Nothing in the step reads workspace memory it has not
// written in the same call, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[VK_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_of_current_device(void** out) {
  int dev = -1;
  VK_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    vk_set_error("vk_ros2_step_noalloc_f64: device %d outside 0..%d", dev, kMaxDevices - 1);
    return VK_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      vk_set_error("vk_ros2_step_noalloc_f64: hipGetSymbolAddress of the scratch array on "
                   "device %d failed: %s", dev, hipGetErrorString(e));
      return VK_ERR_HIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return VK_OK;
}

struct TimingOff {
  TimingOff() { g_timing_suppressed = true; }
  ~TimingOff() { g_timing_suppressed = false; }
};
}  // namespace

extern "C" {

int vk_ros2_step_noalloc_f64(vk_batch b, const double* y, const double* k, const double* atm,
                             const double* dt, const uint8_t* fix_mask, double* sol,
                             double* delta, int32_t* status, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_ros2_step_noalloc_f64")) return rc;
  if (!y || !k || !atm || !dt || !sol || !delta) {
    vk_set_error("vk_ros2_step_noalloc_f64: null pointer");
    return VK_ERR_ARG;
  }
  const size_t need = vk_ros2_workspace_bytes(b);
  if (need > (size_t)VK_NOALLOC_SCRATCH_BYTES) {
    vk_set_error("vk_ros2_step_noalloc_f64: this batch needs %zu bytes of workspace, more than "
                 "the %llu-byte scratch array; split the batch or call vk_ros2_step_f64 with a "
                 "workspace", need, (unsigned long long)VK_NOALLOC_SCRATCH_BYTES);
    return VK_ERR_UNSUPPORTED;
  }
  void* ws = nullptr;
  if (int rc = scratch_of_current_device(&ws)) return rc;
  TimingOff no_events;
  // the buckets run in order on the caller's stream (no helper streams)
  return ros2_step_impl(b, y, k, atm, dt, fix_mask, sol, delta, ws, status, nullptr,
                        (hipStream_t)stream, false);
}

static int check_ss(vk_batch b, const double* atm, const vk_ss_params* p, const vk_ss_state* s) {
  if (int rc = check_batch(b)) return rc;
  if (!atm || !p || !s || !s->y || !s->y_prev || !s->ymix || !s->sol || !s->delta_arr ||
      !s->compo || !s->atom_ini || !s->atom_loss || !s->atom_loss_prev || !s->dt || !s->t ||
      !s->delta || !s->rel || !s->accept || !s->retry || !s->nega || !s->loss || !s->dcount ||
      !s->status || !s->attempts) {
    vk_set_error("steady-state call: null pointer in arguments");
    return VK_ERR_ARG;
  }
  if (p->na < 1 || p->na > VK_MAX_ATOMS) {
    vk_set_error("steady-state call: na=%d outside 1..%d", p->na, VK_MAX_ATOMS);
    return VK_ERR_ARG;
  }
  if ((s->pin == nullptr) != (s->off_pin == nullptr) || p->update_frq < 0) {
    vk_set_error("steady-state call: pin and off_pin must be given together; update_frq >= 0");
    return VK_ERR_ARG;
  }
  if (p->conv_mode != 0 && (p->conv_mode != 1 || p->conv_step < 2 || !s->ring_y || !s->ring_t ||
                            !s->conv || !s->hp || !s->off_hp || !s->conv_ign ||
                            !s->off_ign)) {
    vk_set_error("steady-state call: conv_mode %d needs conv_step >= 2 and the ring and conv "
                 "arrays", p->conv_mode);
    return VK_ERR_ARG;
  }
  return VK_OK;
}

int vk_ss_init_f64(vk_batch b, const double* atm, const vk_ss_params* p, const vk_ss_state* s,
                   void* stream) {
  g_err[0] = 0;
  if (int rc = check_ss(b, atm, p, s)) return rc;
  if (b->slots_mode) {
    vk_set_error("vk_ss_init_f64: a slot batch initialises placed slots with vk_ss_init_range_f64");
    return VK_ERR_ARG;
  }
  return vk_launch_ss_init(b->d_cols, b->ncols, atm, *p, *s, (hipStream_t)stream);
}

int vk_ss_init_range_f64(vk_batch b, int first, int count, const double* atm,
                         const vk_ss_params* p, const vk_ss_state* s, void* stream) {
  g_err[0] = 0;
  if (int rc = check_ss(b, atm, p, s)) return rc;
  if (first < 0 || count < 1 || first + count > b->ncols) {
    vk_set_error("vk_ss_init_range_f64: columns [%d, %d) outside [0, %d)", first, first + count,
                 b->ncols);
    return VK_ERR_ARG;
  }
  return vk_launch_ss_init_range(b->d_cols, b->ncols, first, count, atm, *p, *s,
                                 (hipStream_t)stream);
}

int vk_ss_attempt_f64(vk_batch b, const double* k, const double* atm, const vk_ss_params* p,
                      const vk_ss_state* s, void* ws, size_t ws_bytes, void* stream) {
  g_err[0] = 0;
  if (int rc = check_ss(b, atm, p, s)) return rc;
  if (!k || !ws || ws_bytes < vk_ros2_workspace_bytes(b)) {
    vk_set_error("vk_ss_attempt_f64: missing rates or workspace (%zu < %zu bytes)", ws_bytes,
                 vk_ros2_workspace_bytes(b));
    return VK_ERR_ARG;
  }
  hipStream_t st = (hipStream_t)stream;
  if (int rc = ros2_step_impl(b, s->y, k, atm, s->dt, nullptr, s->sol, s->delta_arr, ws,
                              s->status, s->status, st))
    return rc;
  KernelTimer t(5, st);
  return vk_launch_ss_control(b->d_cols, b->ncols, atm, *p, *s, st);
}

int vk_photo_f64(vk_batch b, int parts, const vk_photo_setup* setup, const vk_photo_state* state,
                 const double* y, double* k, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_photo_f64")) return rc;
  if (!setup || !state || !setup->off_p || !setup->off_l || !setup->bins || !setup->sflux_top ||
      !state->tau || !state->w0 || !state->sflux || !state->dflux_d || !state->dflux_u ||
      !state->aflux || !state->J || ((parts & 1) && (!y || !setup->dz))) {
    vk_set_error("vk_photo_f64: null pointer in arguments");
    return VK_ERR_ARG;
  }
  if (setup->nbin < 2 || setup->n_br < 0 || (setup->n_abs > 0 && !setup->absp_cross) ||
      (setup->n_br > 0 && (!setup->cross_J || !setup->branch_re_idx || !setup->branch_active))) {
    vk_set_error("vk_photo_f64: inconsistent photolysis setup");
    return VK_ERR_ARG;
  }
  return vk_launch_photo(parts, b->d_cols, b->ncols, b->max_nz, *setup, *state, y, k,
                         (hipStream_t)stream);
}

int vk_ss_vjp_f64(vk_batch b, const double* y, const double* k, const double* atm,
                  const double* lam, double* cot_k, double* cot_atm, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_ss_vjp_f64")) return rc;
  if (!y || !k || !atm || !lam || (!cot_k && !cot_atm)) {
    vk_set_error("vk_ss_vjp_f64: null pointer in arguments");
    return VK_ERR_ARG;
  }
  return vk_launch_ss_vjp(b->d_cols, b->ncols, b->max_nz, b->max_ni, b->max_rec, b->d_nets, y,
                          atm, k, lam, cot_k, cot_atm, (hipStream_t)stream);
}

int vk_step_error_f64(vk_batch b, const double* sol, const double* delta, double* out,
                      int64_t stride, void* stream) {
  g_err[0] = 0;
  if (int rc = check_plain(b, "vk_step_error_f64")) return rc;
  if (!sol || !delta || !out || stride < 1) {
    vk_set_error("vk_step_error_f64: null pointer or stride < 1");
    return VK_ERR_ARG;
  }
  return vk_launch_delta_max(b->d_cols, b->ncols, sol, delta, out, stride, (hipStream_t)stream);
}

int vk_tune_set(int knob, int sl, int value) {
  g_err[0] = 0;
  if (knob == VK_TUNE_SPLIT) {
    if (value < 1 || value > 8) {
      vk_set_error("vk_tune_set: VK_TUNE_SPLIT takes 1..8, not %d", value);
      return VK_ERR_ARG;
    }
    g_split = value;
    return VK_OK;
  }
  if (knob == VK_TUNE_REFINE) {
    if (value < 0 || value > 2) {
      vk_set_error("vk_tune_set: VK_TUNE_REFINE takes 0, 1 or 2, not %d", value);
      return VK_ERR_ARG;
    }
    g_refine = value;
    return VK_OK;
  }
  if (knob == VK_TUNE_KDIRECT) {
    if (value != 0 && value != 1) {
      vk_set_error("vk_tune_set: VK_TUNE_KDIRECT takes 0 or 1, not %d", value);
      return VK_ERR_ARG;
    }
    g_k_direct = value;
    return VK_OK;
  }
  if (knob == VK_TUNE_HOLDX) {
    if (value != 0 && value != 1) {
      vk_set_error("vk_tune_set: VK_TUNE_HOLDX takes 0 or 1, not %d", value);
      return VK_ERR_ARG;
    }
    g_solve_hold = value;
    return VK_OK;
  }
  if (knob == VK_TUNE_MOLPRE) {
    if (value != 0 && value != 1) {
      vk_set_error("vk_tune_set: VK_TUNE_MOLPRE takes 0 or 1, not %d", value);
      return VK_ERR_ARG;
    }
    g_mol_precompute = value;
    return VK_OK;
  }
  if (knob != VK_TUNE_FACTOR || sl < 1 || sl > 7) {
    vk_set_error("vk_tune_set: unknown knob %d or bucket %d", knob, sl);
    return VK_ERR_ARG;
  }
  if (value != VK_FACTOR_TILE && !(value == VK_FACTOR_ROWS && sl <= 6) &&
      value != VK_FACTOR_BLOCKED && value != VK_FACTOR_TILE2 && value != VK_FACTOR_WAVE &&
      !(value == VK_FACTOR_TILE8 && sl % 2 == 0) && value != VK_FACTOR_CR) {
    vk_set_error("vk_tune_set: factor variant %d not available for bucket %d", value, sl);
    return VK_ERR_UNSUPPORTED;
  }
  g_factor_variant[sl] = value;
  return VK_OK;
}
int vk_tune_get(int knob, int sl) {
  if (knob == VK_TUNE_MOLPRE) return g_mol_precompute;
  if (knob == VK_TUNE_HOLDX) return g_solve_hold;
  if (knob == VK_TUNE_KDIRECT) return g_k_direct;
  if (knob == VK_TUNE_REFINE) return g_refine;
  if (knob == VK_TUNE_SPLIT) return g_split;
  if (knob != VK_TUNE_FACTOR || sl < 1 || sl > 7) return -1;
  return g_factor_variant[sl];
}

}  // extern "C"

// ---------------------------------------------------------------------------
// rate construction (vk_rates.hip)
// ---------------------------------------------------------------------------

int vk_launch_rates_eval(const VkRatesDev& R, int what, int nz, const double* T, const double* M,
                         int lowt_on, const double* in_a, const double* in_b, double* out,
                         hipStream_t stream);
int vk_launch_atm_refresh(const VkCol* cols, int ncols, int max_nz, int mode, const double* ymix,
                          const double* y, const double* sol, const vk_ss_params& p, double* atm,
                          const vk_refresh_state& r, hipStream_t stream);
int vk_launch_rates_vjp(const VkRatesDev& R, int nz, const double* T, const double* M,
                        int lowt_on, const double* kbar, double* Tbar, double* Mbar,
                        hipStream_t stream);
int vk_launch_rates_build_batch(const VkCol* cols, int ncols, int max_ni, int max_nz,
                                const VkRatesDev* rates, const double* atm, int lowt_on,
                                double* k, hipStream_t stream);

int vk_rates_create(int nr, int ni, int T, int stop_rev_indx, const int32_t* form,
                    const int32_t* lowt, const int32_t* has_rev, const double* par,
                    const int64_t* reactant_idx, const double* reactant_stoich,
                    const int64_t* product_idx, const double* product_stoich,
                    const double* nasa9, const uint8_t* zero_rows, vk_rates* out) {
  g_err[0] = 0;
  if (!out || nr < 0 || ni < 1 || ni > 255 || T < 1 || T > 4 ||
      (nr > 0 && (!form || !lowt || !has_rev || !par || !reactant_idx || !reactant_stoich ||
                  !product_idx || !product_stoich))) {
    vk_set_error("vk_rates_create: bad arguments (nr=%d ni=%d T=%d)", nr, ni, T);
    return VK_ERR_ARG;
  }
  // Row i of the full build is held at zero: the host mask, row 0, rows past
  // stop_rev_indx at odd offsets (gibbs.py:149-150); a reverse row also at
  // or past stop_rev_indx (gibbs.py:141).
  auto zero = [&](int i) {
    if (zero_rows && zero_rows[i]) return true;
    if (i == 0) return true;
    return i >= stop_rev_indx + 1 && ((i - stop_rev_indx - 1) & 1) == 0;
  };
  const int npairs = (nr + 1) / 2;
  std::vector<VkRatePair> pairs(std::max(npairs, 1));
  for (int p = 0; p < npairs; ++p) {
    const int i = 2 * p + 1;
    VkRatePair& q = pairs[p];
    memset(&q, 0, sizeof(q));
    if (form[i] < VK_KF_ZERO || form[i] > VK_KF_TROE_OH_CH3 || lowt[i] < VK_LOWT_NONE ||
        lowt[i] > VK_LOWT_C2H5) {
      vk_set_error("vk_rates_create: row %d has form %d / cap %d", i, form[i], lowt[i]);
      return VK_ERR_ARG;
    }
    q.form = form[i];
    q.lowt = lowt[i];
    for (int k = 0; k < 6; ++k) q.par[k] = par[6 * i + k];
    double sr = 0.0, sp = 0.0;
    for (int s = 0; s < 4; ++s) q.sp[s] = q.sp[4 + s] = (uint8_t)ni;
    for (int s = 0; s < T; ++s) {
      const int64_t e = (int64_t)i * T + s;
      const int64_t ri = reactant_idx[e], pi = product_idx[e];
      const double rs = reactant_stoich[e], ps = product_stoich[e];
      if (ri < 0 || ri > ni || pi < 0 || pi > ni || (double)(float)rs != rs ||
          (double)(float)ps != ps) {
        vk_set_error("vk_rates_create: row %d slot %d: species index out of range or "
                     "stoichiometry not exact in fp32", i, s);
        return VK_ERR_ARG;
      }
      q.sp[s] = (uint8_t)ri;
      q.st[s] = (float)rs;
      q.sp[4 + s] = (uint8_t)pi;
      q.st[4 + s] = (float)ps;
      sr += rs;
    }
    for (int s = 0; s < T; ++s) sp += product_stoich[(int64_t)i * T + s];
    q.dn = sr - sp;
    q.flags = (zero(i) ? VK_PAIR_ZERO_FWD : 0) |
              ((i + 1 > nr || zero(i + 1) || i + 1 >= stop_rev_indx) ? VK_PAIR_ZERO_REV : 0) |
              (has_rev[i] ? VK_PAIR_HAS_REV : 0);
  }
  size_t off = 0;
  auto take = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~size_t(255); return o; };
  const size_t o_p = take(sizeof(VkRatePair) * pairs.size()),
               o_n = nasa9 ? take((size_t)ni * 20 * 8) : 0, o_z = zero_rows ? take(nr + 1) : 0;
  std::vector<char> host(off, 0);
  memcpy(&host[o_p], pairs.data(), sizeof(VkRatePair) * pairs.size());
  if (nasa9) memcpy(&host[o_n], nasa9, (size_t)ni * 20 * 8);
  if (zero_rows) memcpy(&host[o_z], zero_rows, nr + 1);
  vk_rates_s* r = new vk_rates_s();
  if (hipMalloc(&r->buf, off) != hipSuccess ||
      hipMemcpy(r->buf, host.data(), off, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    if (r->buf) (void)hipFree(r->buf);
    delete r;
    vk_set_error("vk_rates_create: device allocation or upload of %zu bytes failed", off);
    return VK_ERR_NOMEM;
  }
  char* b = (char*)r->buf;
  VkRatesDev& d = r->dev;
  d.nr = nr;
  d.ni = ni;
  d.npairs = npairs;
  d.stop_rev = stop_rev_indx;
  d.pairs = (const VkRatePair*)(b + o_p);
  d.nasa9 = nasa9 ? (const double*)(b + o_n) : nullptr;
  d.zero_rows = zero_rows ? (const uint8_t*)(b + o_z) : nullptr;
  *out = r;
  return VK_OK;
}

int vk_rates_destroy(vk_rates r) {
  if (!r) return VK_OK;
  if (r->buf) (void)hipFree(r->buf);
  delete r;
  return VK_OK;
}

int vk_rates_eval_f64(vk_rates r, int what, int nz, const double* T, const double* M,
                      int lowt_on, const double* in_a, const double* in_b, double* out,
                      void* stream) {
  g_err[0] = 0;
  if (!r || nz < 1 || !T || !out || what < VK_RATES_FORWARD || what > VK_RATES_BUILD) {
    vk_set_error("vk_rates_eval_f64: bad arguments (what=%d nz=%d)", what, nz);
    return VK_ERR_ARG;
  }
  const bool needs_m = (what == VK_RATES_FORWARD || what == VK_RATES_CAPS ||
                        what == VK_RATES_BUILD);
  const bool needs_nasa = (what == VK_RATES_GIBBS || what == VK_RATES_BUILD);
  if ((needs_m && !M) || (needs_nasa && !r->dev.nasa9) ||
      ((what == VK_RATES_CAPS || what == VK_RATES_KEQ || what == VK_RATES_FILL) && !in_a) ||
      (what == VK_RATES_FILL && !in_b)) {
    vk_set_error("vk_rates_eval_f64: mode %d needs %s", what,
                 needs_nasa && !r->dev.nasa9 ? "NASA-9 coefficients" : "an input it was not given");
    return VK_ERR_ARG;
  }
  return vk_launch_rates_eval(r->dev, what, nz, T, M, lowt_on, in_a, in_b, out,
                              (hipStream_t)stream);
}

int vk_rates_vjp_f64(vk_rates r, int nz, const double* T, const double* M, int lowt_on,
                     const double* k_bar, double* T_bar, double* M_bar, void* stream) {
  g_err[0] = 0;
  if (!r || nz < 1 || !T || !M || !k_bar || (!T_bar && !M_bar) || !r->dev.nasa9) {
    vk_set_error("vk_rates_vjp_f64: bad arguments (nz=%d)%s", nz,
                 r && !r->dev.nasa9 ? ", handle has no NASA-9 coefficients" : "");
    return VK_ERR_ARG;
  }
  return vk_launch_rates_vjp(r->dev, nz, T, M, lowt_on, k_bar, T_bar, M_bar,
                             (hipStream_t)stream);
}

int vk_batch_set_rates(vk_batch b, const vk_rates* per_net, int n) {
  if (int rc = check_plain(b, "vk_batch_set_rates")) return rc;
  if (!per_net || n != (int)b->net_dims.size()) {
    vk_set_error("vk_batch_set_rates: need one rates handle per batch network (%d), got %d",
                 (int)b->net_dims.size(), n);
    return VK_ERR_ARG;
  }
  std::vector<VkRatesDev> h(n);
  for (int i = 0; i < n; ++i) {
    if (!per_net[i] || !per_net[i]->dev.nasa9 || per_net[i]->dev.ni != b->net_dims[i].first ||
        per_net[i]->dev.nr != b->net_dims[i].second) {
      vk_set_error("vk_batch_set_rates: handle %d does not match network %d (ni=%d nr=%d) or "
                   "has no NASA-9 coefficients", i, i, b->net_dims[i].first,
                   b->net_dims[i].second);
      return VK_ERR_ARG;
    }
    h[i] = per_net[i]->dev;
  }
  if (b->n_rates != n) {
    if (b->d_rates) (void)hipFree(b->d_rates);
    b->d_rates = nullptr;
    b->n_rates = 0;
    if (hipMalloc(&b->d_rates, sizeof(VkRatesDev) * n) != hipSuccess) {
      (void)hipGetLastError();
      vk_set_error("vk_batch_set_rates: device allocation failed");
      return VK_ERR_NOMEM;
    }
    b->n_rates = n;
  }
  VK_CHECK_HIP(hipMemcpy(b->d_rates, h.data(), sizeof(VkRatesDev) * n, hipMemcpyHostToDevice));
  return VK_OK;
}

int vk_rates_build_batch_f64(vk_batch b, int lowt_on, const double* atm, double* k,
                             void* stream) {
  if (int rc = check_plain(b, "vk_rates_build_batch_f64")) return rc;
  if (!atm || !k || !b->d_rates) {
    vk_set_error("vk_rates_build_batch_f64: %s", !b->d_rates ? "no rates bound (vk_batch_set_rates)"
                                                               : "null pointer");
    return VK_ERR_ARG;
  }
  return vk_launch_rates_build_batch(b->d_cols, b->ncols, b->max_ni, b->max_nz, b->d_rates, atm,
                                     lowt_on, k, (hipStream_t)stream);
}

// ---------------------------------------------------------------------------
// atmosphere refresh (vk_refresh.hip)
// ---------------------------------------------------------------------------

int64_t vk_refresh_record_layout(int nz, int ni, int64_t* off) {
  if (nz < 2 || ni < 1) return -1;
  const VkRefOff o = vk_refresh_offsets(nz, ni);
  if (off) {
    off[0] = o.pico;
    off[1] = o.mass;
    off[2] = o.esc;
    off[3] = o.scal;
  }
  return o.total;
}

int vk_atm_refresh_f64(vk_batch b, int mode, const double* ymix, const double* y,
                       const double* sol, const vk_ss_params* p, double* atm,
                       const vk_refresh_state* r, void* stream) {
  g_err[0] = 0;
  if (!b || mode < 1 || mode > 3 || !atm || !r || !r->rec || !r->off_r || !r->geo ||
      !r->off_g) {
    vk_set_error("vk_atm_refresh_f64: bad arguments (mode=%d)", mode);
    return VK_ERR_ARG;
  }
  if (sol ? (!ymix || !p) : (((mode & 1) && !ymix) || ((mode & 2) && !y))) {
    vk_set_error("vk_atm_refresh_f64: %s", sol ? "the in-loop form needs ymix and the params"
                                               : "mode needs ymix (bit 1) / y (bit 2)");
    return VK_ERR_ARG;
  }
  vk_ss_params pp{};
  if (p) pp = *p;
  return vk_launch_atm_refresh(b->d_cols, b->ncols, b->max_nz, mode, ymix, y, sol, pp, atm, *r,
                               (hipStream_t)stream);
}
