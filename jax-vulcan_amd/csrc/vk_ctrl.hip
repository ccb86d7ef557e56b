// Per-column adaptive step control, kept on the device.
//
// Restates the accept/reject part of outer_loop.body_fn:
//   clip (outer_loop.py:196-224), atom loss (:171-178), aggregated delta
//   (:227-250), accept / force-accept / reject and the counters (:741-768),
//   hydrostatic rebalance n0 * ymix (:847-861), state update (:900-917) and
//   the dt update (:181-193, 1031-1038),
// plus the stopping rule of the plain steady-state run: after an accepted
// step, the largest masked relative change |y - y_prev| / y (the quantity of
// _f_dy, outer_loop.py:2163-2184) below conv_tol once t > trun_min and
// accept_count > count_min (the "ready" gate of cond_fn, :706-711), or
// count_max accepted steps.
//
// One 1024-thread workgroup per column, one layer per wave at a time, with
// fixed reduction trees, so every column's result is independent of the
// batch.
#include "vk_common.h"

namespace {

constexpr int NT = VK_NT;

__device__ inline double clip_val(double v, double ymix_old, const vk_ss_params& p) {
  return vk_clip_val(v, ymix_old, p);
}

// Control kernel workgroup: 16 waves, one layer per wave at a time.
constexpr int NTC = 1024;

template <int NTB = NT, typename T, typename Op>
__device__ inline T block_reduce(T v, T* red, Op op) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T r = red[0];
  for (int w = 1; w < NTB / 64; ++w) r = op(r, red[w]);
  return r;
}

struct AddD { __device__ double operator()(double a, double b) const { return a + b; } };
struct MaxD { __device__ double operator()(double a, double b) const { return a > b ? a : b; } };
// jnp.max semantics: a NaN operand makes the maximum NaN
__device__ inline double nan_max(double a, double b) { return (a > b || a != a) ? a : b; }
struct NanMaxD { __device__ double operator()(double a, double b) const { return nan_max(a, b); } };
// jnp.clip semantics (maximum then minimum, both NaN-propagating)
__device__ inline double nan_clip(double x, double lo, double hi) {
  return x != x ? x : fmin(fmax(x, lo), hi);
}
struct OrI { __device__ int operator()(int a, int b) const { return a | b; } };

// Sum over the 64 lanes of a wave by a fixed butterfly; every lane returns
// lane 0's total, so the result is one rounding sequence for the wave.
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return __shfl(v, 0, 64);
}

}  // namespace

__global__ __launch_bounds__(NTC) void k_ss_control(const VkCol* __restrict__ cols, int ncols,
                                                    const double* __restrict__ atm,
                                                    vk_ss_params p, vk_ss_state s) {
  __shared__ double redd[NTC / 64];
  __shared__ int redi[NTC / 64];
  const int c = blockIdx.x;
  if (c >= ncols) return;
  if (s.status[c] & VK_ST_DONE) return;
  const VkCol col = cols[c];
  const int nz = col.nz, ni = col.ni, tid = threadIdx.x;
  const int na = p.na;
  const VkAtmOff ao = vk_atm_offsets(nz, ni);
  const double* n0 = atm + col.off_a + ao.M;
  const double* gas = atm + col.off_a + ao.gas;
  double* y = s.y + col.off_s;
  double* yp = s.y_prev + col.off_s;
  double* ymix = s.ymix + col.off_s;
  const double* sol = s.sol + col.off_s;
  const double* dar = s.delta_arr + col.off_s;
  const double* compo = s.compo + col.off_s;  // (ni, na)

  // Every pass takes one layer per wave (layers wave, wave + 4, ...) and
  // species lane, lane + 64 of it, so loads are coalesced along the row and a
  // layer's sums are one fixed shuffle tree; the same thread owns the same
  // (layer, species) cells in every pass.
  const int lane = tid & 63, wave = tid >> 6;
  constexpr int NW = NTC / 64;
  constexpr int PER = 2;   // species per lane: ni <= 128

  // ---- pass 1: clip, delta, sign flags, atom sums ---------------------------
  double dmax = 0.0;
  // 1: a value after the clip that is not >= 0 (NaN included: all(sol_clip >= 0)
  //    fails on NaN, outer_loop.py:747), 2: a non-finite value,
  // 4: a value after the clip that is < 0 (any(sol_clip < 0), :757)
  int flags = 0;
  double at_clip[VK_MAX_ATOMS];
  for (int a = 0; a < VK_MAX_ATOMS; ++a) at_clip[a] = 0.0;
  for (int z = wave; z < nz; z += NW) {
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = lane + 64 * u;
      if (i >= ni) continue;
      const int64_t e = (int64_t)z * ni + i;
      const double sv = sol[e], mo = ymix[e];
      const double sc = clip_val(sv, mo, p);
      if (!(sc >= 0.0)) flags |= 1;
      if (sc < 0.0) flags |= 4;
      if (!isfinite(sv)) flags |= 2;
      for (int a = 0; a < na; ++a) at_clip[a] += compo[i * na + a] * sc;
      // aggregated delta (outer_loop.py:238-248) with the pre-step ymix
      double m = (mo < p.mtol) ? 0.0 : dar[e];
      if (sc < p.atol) m = 0.0;
      if (p.zero_bot_row && z == 0) m = 0.0;
      if (sc > 0.0) dmax = nan_max(dmax, m / fmax(fabs(sc), 1e-300));
    }
  }
  dmax = block_reduce<NTC>(dmax, redd, NanMaxD());
  flags = block_reduce<NTC>(flags, redi, OrI());
  for (int a = 0; a < na; ++a) at_clip[a] = block_reduce<NTC>(at_clip[a], redd, AddD());

  // ---- accept / reject (outer_loop.py:741-768) -------------------------------
  double* al = s.atom_loss + (int64_t)c * VK_MAX_ATOMS;
  double* alp = s.atom_loss_prev + (int64_t)c * VK_MAX_ATOMS;
  const double* ai = s.atom_ini + (int64_t)c * VK_MAX_ATOMS;
  double loss_diff = 0.0;
  for (int a = 0; a < na; ++a) {
    const double aln = (at_clip[a] - ai[a]) / ai[a];
    loss_diff = nan_max(loss_diff, fabs(aln - alp[a]));
  }
  const double dt = s.dt[c];
  const int retry = s.retry[c];
  const bool all_nonneg = !(flags & 1);
  const bool accept = all_nonneg && (loss_diff < p.loss_eps) && (dmax <= p.rtol);
  const double next_dt_rej = dt * p.dt_var_min;
  const bool force = ((next_dt_rej < p.dt_min) || (retry >= p.batch_max_retries)) && !accept;
  const bool do_accept = accept || force;

  // ---- pass 2: state update -----------------------------------------------
  // accepted: y = n0 * ymix_new (rebalanced), ymix = y / ysum(y)
  // forced:   y = max(y_prev, 0), ymix unchanged
  // rejected: y = y_prev, ymix unchanged
  double relmax = 0.0;
  double at_b[VK_MAX_ATOMS];
  for (int a = 0; a < VK_MAX_ATOMS; ++a) at_b[a] = 0.0;
  // per-column constraints on accepted states (outer_loop.py:855-869)
  double* pin = s.pin ? s.pin + s.off_pin[c] : nullptr;
  const int pflags = pin ? (int)pin[0] : 0;
  const double t_old = s.t[c];
  for (int z = wave; z < nz; z += NW) {
    // H2/He pin (outer_loop.py:877-882): the snapshot trips on any attempt
    // with t > pin_time, from the bottom layer's pre-step ymix (every lane
    // forms the same values; lane 0 stores them)
    bool h2he_apply = false;
    double h2m = 0.0, hem = 0.0;
    int h2 = -1, he = -1;
    if (z == 0 && (pflags & VK_PIN_H2HE)) {
      double* hp = pin + 2 + 3 * ni;
      h2 = (int)hp[0];
      he = (int)hp[1];
      if (h2 >= 0 && h2 < ni && he >= 0 && he < ni) {
        const bool pinned = hp[3] != 0.0;
        const bool trip = !pinned && t_old > hp[2];
        h2m = trip ? ymix[h2] : hp[4];
        hem = trip ? ymix[he] : hp[5];
        if (lane == 0) {
          hp[3] = (pinned || trip) ? 1.0 : 0.0;
          hp[4] = h2m;
          hp[5] = hem;
        }
        h2he_apply = pinned || trip;
      }
    }
    double sc[PER], yv[PER];
    double ys = 0.0;
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = lane + 64 * u;
      sc[u] = 0.0;
      if (i >= ni) continue;
      const int64_t e = (int64_t)z * ni + i;
      sc[u] = clip_val(sol[e], ymix[e], p);
      if (gas[i] != 0.0) ys += sc[u];
    }
    double ysb = 0.0;
    if (accept) {
      ys = wave_sum(ys);
      // hydrostatic balance of the gas species (outer_loop.py:847-853); the
      // other species keep their clipped value, then the constraints act on
      // the balanced layer
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int i = lane + 64 * u;
        yv[u] = (i < ni && gas[i] != 0.0) ? n0[z] * (sc[u] / ys) : sc[u];
      }
      if (pflags) {
        if (pflags & VK_PIN_CHARGE) {   // e = -sum_i charge_i y_i (outer_loop.py:855-860)
          const int eidx = (int)pin[1];
          double q = 0.0;
#pragma unroll
          for (int u = 0; u < PER; ++u) {
            const int i = lane + 64 * u;
            if (i < ni) q = q + yv[u] * pin[2 + i];
          }
          q = wave_sum(q);
#pragma unroll
          for (int u = 0; u < PER; ++u)
            if (lane + 64 * u == eidx) yv[u] = -q;
        }
#pragma unroll
        for (int u = 0; u < PER; ++u) {
          const int i = lane + 64 * u;
          if (i >= ni || z != 0) continue;
          if (pflags & VK_PIN_ALL_BOT) yv[u] = pin[2 + ni + i];    // outer_loop.py:864-865
          if (pflags & VK_PIN_SP_BOT) {                             // outer_loop.py:866-869
            const double m = pin[2 + 2 * ni + i];
            if (m == m) yv[u] = m * n0[0];
          }
          if (h2he_apply) {                                         // outer_loop.py:884-892
            if (i == h2) yv[u] = h2m * n0[0];
            if (i == he) yv[u] = hem * n0[0];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int i = lane + 64 * u;
        if (i < ni && gas[i] != 0.0) ysb += yv[u];
      }
      ysb = wave_sum(ysb);
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int i = lane + 64 * u;
      if (i >= ni) continue;
      const int64_t e = (int64_t)z * ni + i;
      const double yprev = yp[e];
      double ynew, mixnew = ymix[e];
      if (force) {
        ynew = yprev < 0.0 ? 0.0 : yprev;
      } else if (accept) {
        ynew = yv[u];
        mixnew = ynew / ysb;
        for (int a = 0; a < na; ++a) at_b[a] += compo[i * na + a] * ynew;
      } else {
        ynew = yprev;
      }
      if (do_accept) {
        // relative change vs the previous accepted state (outer_loop.py:2174-2183)
        double dy = fabs(ynew - yprev);
        if (mixnew < p.mtol) dy = 0.0;
        if (ynew < p.atol) dy = 0.0;
        if (ynew > 0.0) {
          const double r = dy / ynew;
          relmax = r > relmax ? r : relmax;
        }
        yp[e] = ynew;
      }
      y[e] = ynew;
      ymix[e] = mixnew;
    }
  }
  relmax = block_reduce<NTC>(relmax, redd, MaxD());
  if (accept)
    for (int a = 0; a < na; ++a) at_b[a] = block_reduce<NTC>(at_b[a], redd, AddD());

  // ---- pass 3 (conv_mode 1): the reference's longdy (outer_loop.py:636-670)
  __shared__ double cv_d[NTC / 64];
  __shared__ int cv_i[NTC / 64];
  double longdy = 0.0, slope_min = 0.0;
  int indx = 0, slot = 0;
  const bool ring = p.conv_mode == 1 && do_accept;
  if (ring) {
    const double t_new = s.t[c] + (force ? p.dt_min : dt);
    const int acc_next = s.accept[c] + 1;
    slot = (acc_next - 1) % p.conv_step;
    double* rt = s.ring_t + (int64_t)c * p.conv_step;
    // the ring entry nearest t * st_factor, never the one this step fills
    const double target = t_new * p.st_factor;
    double bd = INFINITY;
    int bi = 0x7fffffff;
    for (int i = tid; i < p.conv_step; i += NTC) {
      const double d = (i == slot) ? INFINITY : fabs(rt[i] - target);
      if (d < bd || (d == bd && i < bi)) { bd = d; bi = i; }
    }
    // (distance, index) minimum, smallest index on ties: per wave, then waves
    for (int off = 32; off > 0; off >>= 1) {
      const double od = __shfl_xor(bd, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (od < bd || (od == bd && oi < bi)) { bd = od; bi = oi; }
    }
    if (lane == 0) {
      cv_d[wave] = bd;
      cv_i[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      double d0 = cv_d[0];
      int i0 = cv_i[0];
      for (int w = 1; w < NW; ++w)
        if (cv_d[w] < d0 || (cv_d[w] == d0 && cv_i[w] < i0)) { d0 = cv_d[w]; i0 = cv_i[w]; }
      cv_i[0] = (i0 == 0x7fffffff) ? 0 : i0;
    }
    __syncthreads();
    indx = cv_i[0];
    const double* Hp = s.hp + s.off_hp[c];
    const double* ign = s.conv_ign + s.off_ign[c];
    const double* Kzz = atm + col.off_a + ao.Kzz;
    double* ry = s.ring_y + (int64_t)p.conv_step * col.off_s;
    const double* yold = ry + (int64_t)indx * nz * ni;
    double* ynew_slot = ry + (int64_t)slot * nz * ni;
    double lmax = 0.0, smin = INFINITY;
    for (int z = wave; z < nz; z += NW) {
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        const int i = lane + 64 * u;
        if (i >= ni) continue;
        const int64_t e = (int64_t)z * ni + i;
        const double yv = y[e], mv = ymix[e];
        double ld = fabs((yv - yold[e]) / n0[z]);
        if (mv < p.mtol_conv || yv < p.atol || ign[i] != 0.0) ld = 0.0;
        const double r = (mv > 0.0) ? ld / fmax(mv, 1e-300) : 0.0;
        lmax = r > lmax ? r : lmax;
        ynew_slot[e] = yv;
      }
      if (lane == 0 && z + 1 < nz) {
        const double h = 0.1 * Hp[z];
        const double q = Kzz[z] / (h * h);
        smin = q < smin ? q : smin;
      }
    }
    longdy = block_reduce<NTC>(lmax, redd, MaxD());
    smin = -block_reduce<NTC>(-smin, redd, MaxD());
    slope_min = fmax(fmin(smin, 1e-8), 1e-10);   // outer_loop.py:679-683
  }

  if (tid == 0) {
    // atom loss bookkeeping (outer_loop.py:898, 914-917)
    if (accept) {
      for (int a = 0; a < na; ++a) {
        al[a] = (at_b[a] - ai[a]) / ai[a];
        alp[a] = al[a];
      }
    } else if (force) {
      for (int a = 0; a < na; ++a) al[a] = alp[a];
    }
    const bool delta_too_big = dmax > p.rtol;
    const bool reject = !do_accept;
    if (reject && delta_too_big) s.dcount[c] += 1;
    if (reject && !delta_too_big && (flags & 4)) s.nega[c] += 1;
    if (reject && !delta_too_big && !(flags & 4)) s.loss[c] += 1;
    const double t_old = s.t[c];
    const double t_new = do_accept ? t_old + (force ? p.dt_min : dt) : t_old;
    s.t[c] = t_new;
    s.delta[c] = dmax;
    const int acc = s.accept[c] + (do_accept ? 1 : 0);
    s.accept[c] = acc;
    s.retry[c] = do_accept ? 0 : retry + 1;
    s.attempts[c] += 1;
    // dt update (outer_loop.py:181-193, 1031-1038)
    const double base = force ? p.dt_min : dt;
    double dt_new;
    if (do_accept) {
      const double de = (dmax == 0.0) ? 0.01 * p.rtol : dmax;
      const double h = nan_clip(0.9 * sqrt(p.rtol / de), p.dt_var_min, p.dt_var_max);
      dt_new = nan_clip(base * h, p.dt_min, p.dt_max);
    } else {
      dt_new = next_dt_rej;
    }
    s.dt[c] = dt_new;
    int st = s.status[c];
    const bool flux_ok = !s.aflux_change || s.aflux_change[c] < p.flux_cri;
    if (p.conv_mode == 1) {
      double* cv = s.conv + (int64_t)c * 4;
      if (ring) {
        double* rt = s.ring_t + (int64_t)c * p.conv_step;
        const double ldt = longdy / fmax(t_new - rt[indx], 1e-300);
        rt[slot] = t_new;
        cv[0] = longdy;
        cv[1] = ldt;
        // stall bookkeeping: a 5 % drop is a new minimum (outer_loop.py:986-1003)
        if (longdy < cv[2] * 0.95) {
          cv[2] = longdy;
          cv[3] = 0.0;
        } else {
          cv[3] += 1.0;
        }
        s.rel[c] = longdy;
      }
      // cond_fn (outer_loop.py:672-713), on the carried longdy / longdydt
      const double ld = cv[0], ldt = cv[1];
      bool conv = ((ld < p.yconv_cri) && (ldt < p.slope_cri)) ||
                  ((ld < p.yconv_min) && (ldt < slope_min));
      conv = conv && flux_ok;
      const bool stalled = (cv[3] > (double)p.conv_stall_window) && (cv[2] < p.yconv_min) &&
                           (ld < p.yconv_min) && flux_ok;
      const bool ready = (t_new > p.trun_min) && (acc > p.count_min);
      if (do_accept && ready && (conv || stalled)) st |= VK_ST_CONVERGED;
      if (acc > p.count_max || t_new > p.runtime) st |= VK_ST_MAXSTEPS;
    } else {
      if (do_accept) {
        s.rel[c] = relmax;
        const bool ready = (t_new > p.trun_min) && (acc > p.count_min);
        if (ready && relmax < p.conv_tol && flux_ok) st |= VK_ST_CONVERGED;
      }
      if (acc >= p.count_max) st |= VK_ST_MAXSTEPS;
    }
    if (flags & 2) st |= VK_ST_NONFINITE;
    s.status[c] = st;
    // atmosphere refresh after this attempt (outer_loop.py:813-820): accepted
    // or forced, counted before the increment
    if (s.refresh_due)
      s.refresh_due[c] = (do_accept && p.update_frq > 0 &&
                          ((acc - 1) % p.update_frq) == 0) ? 1 : 0;
    // photolysis cadence: the ini -> final switch on the carried longdy and
    // longdydt (outer_loop.py:1040-1049); under conv_mode 0 the last
    // accepted step's relative change stands in for longdy
    int frq = s.photo_cad ? s.photo_cad[c] : p.photo_frq;
    if (s.photo_cad && p.photo_frq_final > 0 && frq != p.photo_frq_final) {
      double ld, ldt;
      if (p.conv_mode == 1) {
        ld = s.conv[(int64_t)c * 4];
        ldt = s.conv[(int64_t)c * 4 + 1];
      } else {
        ld = s.rel[c];
        ldt = s.rel[c] / dt;
      }
      if (ld < p.photo_sw_longdy && ldt < p.photo_sw_longdydt) {
        frq = p.photo_frq_final;
        s.photo_cad[c] = frq;
      }
    }
    // photolysis due before the next attempt (outer_loop.py:253-257, 716-724)
    if (s.photo_due) s.photo_due[c] = (s.retry[c] == 0 && frq > 0 && (acc % frq) == 0) ? 1 : 0;
  }
}

// ymix, atom_ini / atom_loss, counters; y and dt are set by the caller.
__global__ __launch_bounds__(NT) void k_ss_init(const VkCol* __restrict__ cols, int ncols,
                                                 const double* __restrict__ atm, vk_ss_params p,
                                                 vk_ss_state s, int first) {
  __shared__ double redd[NT / 64];
  const int c = first + (int)blockIdx.x;
  if (c >= ncols) return;
  const VkCol col = cols[c];
  const int nz = col.nz, ni = col.ni, tid = threadIdx.x, na = p.na;
  const VkAtmOff ao = vk_atm_offsets(nz, ni);
  const double* gas = atm + col.off_a + ao.gas;
  double* y = s.y + col.off_s;
  double* yp = s.y_prev + col.off_s;
  double* ymix = s.ymix + col.off_s;
  const double* compo = s.compo + col.off_s;
  double at[VK_MAX_ATOMS];
  for (int a = 0; a < VK_MAX_ATOMS; ++a) at[a] = 0.0;
  for (int z = tid; z < nz; z += NT) {
    double ys = 0.0;
    for (int i = 0; i < ni; ++i)
      if (gas[i] != 0.0) ys += y[(int64_t)z * ni + i];
    for (int i = 0; i < ni; ++i) {
      const int64_t e = (int64_t)z * ni + i;
      ymix[e] = y[e] / ys;
      yp[e] = y[e];
      for (int a = 0; a < na; ++a) at[a] += compo[i * na + a] * y[e];
    }
  }
  for (int a = 0; a < na; ++a) at[a] = block_reduce(at[a], redd, AddD());
  if (p.conv_mode == 1) {
    // empty ring and longdy state (outer_loop.py:1777-1786)
    double* ry = s.ring_y + (int64_t)p.conv_step * col.off_s;
    const int64_t n = (int64_t)p.conv_step * nz * ni;
    for (int64_t e = tid; e < n; e += NT) ry[e] = 0.0;
    for (int i = tid; i < p.conv_step; i += NT) s.ring_t[(int64_t)c * p.conv_step + i] = 0.0;
    if (tid == 0) {
      // longdy and longdydt start at 1 (state.py:706-707)
      double* cv = s.conv + (int64_t)c * 4;
      cv[0] = 1.0;
      cv[1] = 1.0;
      cv[2] = INFINITY;
      cv[3] = 0.0;
    }
  }
  if (tid == 0) {
    for (int a = 0; a < VK_MAX_ATOMS; ++a) {
      const double v = a < na ? at[a] : 1.0;
      s.atom_ini[(int64_t)c * VK_MAX_ATOMS + a] = v;
      s.atom_loss[(int64_t)c * VK_MAX_ATOMS + a] = 0.0;
      s.atom_loss_prev[(int64_t)c * VK_MAX_ATOMS + a] = 0.0;
    }
    s.t[c] = 0.0;
    s.delta[c] = 0.0;
    s.rel[c] = 1.0;   // the relative change before any accepted step, as longdy (state.py:706)
    s.accept[c] = 0;
    s.retry[c] = 0;
    s.nega[c] = 0;
    s.loss[c] = 0;
    s.dcount[c] = 0;
    s.attempts[c] = 0;
    s.status[c] = 0;
    if (s.photo_due) s.photo_due[c] = p.photo_frq > 0 ? 1 : 0;
    if (s.photo_cad) s.photo_cad[c] = p.photo_frq;
  }
}

int vk_launch_ss_control(const VkCol* cols, int ncols, const double* atm, const vk_ss_params& p,
                         const vk_ss_state& s, hipStream_t st) {
  hipLaunchKernelGGL(k_ss_control, dim3(ncols), dim3(NTC), 0, st, cols, ncols, atm, p, s);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_ss_init(const VkCol* cols, int ncols, const double* atm, const vk_ss_params& p,
                      const vk_ss_state& s, hipStream_t st) {
  hipLaunchKernelGGL(k_ss_init, dim3(ncols), dim3(NT), 0, st, cols, ncols, atm, p, s, 0);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}

int vk_launch_ss_init_range(const VkCol* cols, int ncols, int first, int count,
                            const double* atm, const vk_ss_params& p, const vk_ss_state& s,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_ss_init, dim3(count), dim3(NT), 0, st, cols, ncols, atm, p, s, first);
  VK_CHECK_HIP(hipGetLastError());
  return VK_OK;
}
