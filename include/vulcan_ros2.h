/* C ABI of the gfx950 Ros2 column-step library (libvulcan_hip.so).
 *
 * Every compute entry point takes caller-owned DEVICE buffers (plain
 * pointers), an explicit hipStream_t (passed as void*), runs asynchronously
 * on that stream and returns an int status: VK_OK or an error code whose
 * text is available from vk_last_error() (thread-local).  Numerical trouble
 * (zero pivot, non-finite state) never aborts a batch: it sets bits in a
 * per-column int32 status word instead.
 *
 * Batches.  A vk_batch describes ncols columns packed into flat buffers.
 * Column c has nz[c] layers, ni[c] species and S[c] rate rows, and uses
 * network net[c].  Per-column arrays sit at these offsets (in doubles):
 *   state-shaped (nz, ni)      : off_s[c]   (y, sol, delta, rhs, sup/sub...)
 *   natural blocks (nz, ni, ni): off_s[c]*ni[c]
 *   rate table (S, nz)         : off_k[c]   (the reference's k_arr layout)
 *   atmosphere record          : off_a[c]   (layout: vk_atm_record_layout)
 *   block stacks               : off_b[c]   (nz * vk_block_doubles(ni))
 * vk_batch_uniform() builds the table for B identical columns stacked
 * along a leading axis, which is exactly jax.vmap's layout.
 *
 * Reference interfaces replaced (file:line in imalsky/jax-vulcan):
 *   vk_network_create    chem.to_jax / NetworkArrays           chem.py:28-80
 *   vk_chem_rhs_f64      chem.chem_rhs                         chem.py:83-131
 *   vk_chem_jac_f64      chem.chem_jac_analytical              chem.py:144-214
 *   vk_bt_factor_f64     solver.factor_block_thomas_diag_offdiag solver.py:33-64
 *   vk_bt_solve_f64      solver.solve_block_thomas_diag_offdiag  solver.py:67-112
 *   vk_bt_dense_f64      solver.block_thomas                   solver.py:129-176
 *   vk_diff_grav_f64     jax_step.compute_diff_grav            jax_step.py:67-147
 *   vk_ros2_step_f64     jax_step.jax_ros2_step                jax_step.py:249-315
 *   vk_step_error_f64    integrate.jax_integrate_fixed_dt's delta  integrate.py:30-33
 *   vk_residual_f64      steady_state_grad.steady_state_residual steady_state_grad.py:116-150
 *   vk_ss_vjp_f64        steady_state_grad._ssi_bwd / _ss_bwd parameter VJP
 *                                                               steady_state_grad.py:214-233, 306-327
 *   vk_jacobian_blocks_f64 steady_state_grad._build_jacobian_blocks
 *                                                               steady_state_grad.py:153-175
 */
#ifndef VULCAN_ROS2_H
#define VULCAN_ROS2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  VK_OK = 0,
  VK_ERR_ARG = 1,         /* bad argument (shape, null pointer, range) */
  VK_ERR_HIP = 2,         /* HIP runtime error */
  VK_ERR_NOMEM = 3,       /* device or host allocation failed */
  VK_ERR_UNSUPPORTED = 4  /* size outside what the kernels support */
};

typedef struct vk_network_s* vk_network;
typedef struct vk_batch_s* vk_batch;

/* Last error text of the calling thread ("" when none). */
const char* vk_last_error(void);
/* Library ABI version (major*100 + minor). */
int vk_version(void);
/* Largest ni handled by the block kernels (16 * 7 = 112). */
int vk_max_species(void);

/* Upload one reaction network given in the reference's padded table layout
 * (network.py:366-451): rows 1..nr (row 0 unused), T = max_terms slots per
 * row, PAD = ni marks an empty slot, host pointers.  Builds the species-major
 * production/loss lists and the entry-major Jacobian lists on the host.
 * Replaces chem.to_jax (chem.py:70-80). */
int vk_network_create(int ni, int nr, int T, const int64_t* reactant_idx,
                      const int64_t* product_idx, const double* reactant_stoich,
                      const double* product_stoich, const uint8_t* is_three_body,
                      vk_network* out);
int vk_network_destroy(vk_network net);

/* Batch tables (host arrays of length ncols, copied to the device). */
int vk_batch_create(int ncols, const int32_t* nz, const int32_t* ni, const int32_t* S,
                    const int32_t* net, const int64_t* off_s, const int64_t* off_k,
                    const int64_t* off_a, const int64_t* off_b, const vk_network* nets,
                    int nnets, vk_batch* out);
int vk_batch_uniform(int ncols, int nz, vk_network net, vk_batch* out);
int vk_batch_destroy(vk_batch b);

/* Slot batch for continuous batching (no counterpart in the reference,
 * whose runs are one column or a vmapped uniform batch): nslots columns of
 * fixed capacity -- up to cap_nz layers, cap_S rate rows, cap_T reactant
 * slots, ni <= cap_ni in the same 16-species bucket as cap_ni.  Slot c's
 * arrays sit at off_s = c*cap_nz*cap_ni, off_k = c*cap_S*cap_nz, off_a =
 * c*(record length of (cap_nz, cap_ni)), off_b = c*cap_nz*vk_block_doubles.
 * Slots start empty: the caller keeps their status words at a DONE value
 * until vk_batch_set_column places a column, and only the steady-state
 * entry points accept a slot batch. */
int vk_batch_slots(int nslots, int cap_nz, int cap_ni, int cap_S, int cap_T, vk_batch* out);
/* Place a column of nz layers with network net in a slot: the descriptor is
 * rewritten by a one-thread kernel on the stream, so it is ordered with the
 * launches around it and legal between graph replays. */
int vk_batch_set_column(vk_batch b, int slot, int nz, vk_network net, void* stream);
/* Latency mode (no counterpart in the reference): on = 1 makes every bucket
 * of b use the cyclic-reduction block factor (VK_FACTOR_CR) in
 * vk_ros2_step_f64 and vk_ss_attempt_f64, whatever the process-wide
 * VK_TUNE_FACTOR says; on = 0 follows it again.  For batches of a few
 * columns per bucket, such as serving pools.  Set it before sizing the
 * workspace: vk_ros2_workspace_bytes(b) then includes the 7 block stacks
 * of the reduction levels. */
int vk_batch_set_latency_mode(vk_batch b, int on);
/* Doubles needed by a block stack of one layer with ni species. */
int64_t vk_block_doubles(int ni);
/* Total doubles of the packed block stacks of a batch. */
int64_t vk_batch_block_doubles(vk_batch b);
/* Length (doubles) of one atmosphere record and the offsets of its fields,
 * in this order: Kzz dzi vz Hpi Ti (nz-1 each), Tco g M (nz), ms alpha
 * top_flux bot_flux bot_vdep gas_mask (ni), Dzz (nz-1,ni), vm (nz,ni),
 * vs (nz-1,ni), flags[4] = use_vm_mol use_settling use_topflux use_botflux.
 * Field set and meaning: jax_step.AtmStatic (jax_step.py:26-49). */
int64_t vk_atm_record_layout(int nz, int ni, int64_t* offsets18);

/* out (off_s) = chemistry rhs at y.  M: per-layer (ncols, nz) array for a
 * uniform batch, or NULL to read M from the atmosphere records. */
int vk_chem_rhs_f64(vk_batch b, const double* y, const double* M, const double* atm,
                    const double* k, double* out, void* stream);
/* jac (natural blocks, off_s*ni) = analytical chemistry Jacobian at y. */
int vk_chem_jac_f64(vk_batch b, const double* y, const double* M, const double* atm,
                    const double* k, double* jac, void* stream);

/* Block-Thomas factor.  diag: natural blocks (off_s*ni); sup_d, sub_d:
 * (nz-1, ni) at off_s.  transpose != 0 factors the system whose blocks are
 * diag[j]^T (the adjoint's transposed solve; pass sup/sub swapped as the
 * reference does, steady_state_grad.py:316-317).  inv receives the stack of
 * inverses inv(A'_j) in tile layout (off_b).  status: int32 per column. */
int vk_bt_factor_f64(vk_batch b, const double* diag, const double* sup_d, const double* sub_d,
                     int transpose, double* inv, int32_t* status, void* stream);
/* x = solve with a factored stack; rhs and x at off_s; scratch: state-sized. */
int vk_bt_solve_f64(vk_batch b, const double* inv, const double* sup_d, const double* sub_d,
                    const double* rhs, double* x, double* scratch, void* stream);
/* jax_step.compute_diff_grav (jax_step.py:67-147): the y-independent parts
 * of the molecular-diffusion coefficients from the atmosphere records.
 * A, B, C are (nz, ni) per column at off_s: rows 1..nz-2 are A/B/C_grav_int,
 * row 0 holds bdry0_grav (A, B) and row nz-1 bdry_top_grav (A, C); the
 * entries the reference does not define (C of row 0, B of row nz-1) are 0.
 * dz_ave = 0.5 (dzi[j-1] + dzi[j]) is left to the caller. */
int vk_diff_grav_f64(vk_batch b, const double* atm, double* A, double* B, double* C,
                     void* stream);
/* Dense block-tridiagonal Thomas: solver.block_thomas (solver.py:129-176),
 * full ni x ni off-diagonal blocks, one right-hand side, factor and both
 * sweeps in one launch (one workgroup per column).  diag, sup and sub are
 * natural (ni, ni) blocks at off_s * ni, nz blocks per column (sup[j] = B_j
 * couples layer j to j+1, sub[j] = C_{j+1} couples j+1 to j; block nz-1 of
 * sup and sub is not read); rhs and x at off_s.  inv receives the inverses
 * of the Schur complements (tile layout, off_b); scratch is state-sized;
 * status as vk_bt_factor_f64. */
int vk_bt_dense_f64(vk_batch b, const double* diag, const double* sup, const double* sub,
                    const double* rhs, double* x, double* inv, double* scratch, int32_t* status,
                    void* stream);
/* Latency mode of the two above (no counterpart in the reference; same
 * system as solver.factor_block_thomas_diag_offdiag /
 * solve_block_thomas_diag_offdiag, solver.py:33-112): block cyclic
 * reduction, about log2(nz) levels of one workgroup per block row, for a
 * few columns (the Newton polish and the adjoint solve of one column).
 * Same arguments; the factored levels live in ws (vk_bt_cr_workspace_bytes:
 * 8 block stacks), which the solve reads.  Results agree with the tile
 * factor to the system's conditioning, not bit for bit. */
size_t vk_bt_cr_workspace_bytes(vk_batch b);
int vk_bt_cr_factor_f64(vk_batch b, const double* diag, const double* sup_d, const double* sub_d,
                        int transpose, void* ws, size_t ws_bytes, int32_t* status, void* stream);
int vk_bt_cr_solve_f64(vk_batch b, const double* sup_d, const double* sub_d, const double* rhs,
                       double* x, void* ws, size_t ws_bytes, void* stream);
/* Copy a tile-layout inverse stack into natural (nz, ni, ni) blocks. */
int vk_bt_inverse_natural_f64(vk_batch b, const double* inv, double* nat, void* stream);

/* One Ros2 step for every column: sol, delta at off_s.  dt: per-column
 * device array.  fix_mask: optional uint8 state-shaped mask (NULL for none).
 * ws: device workspace of vk_ros2_workspace_bytes() bytes. */
size_t vk_ros2_workspace_bytes(vk_batch b);
int vk_ros2_step_f64(vk_batch b, const double* y, const double* k, const double* atm,
                     const double* dt, const uint8_t* fix_mask, double* sol, double* delta,
                     void* ws, size_t ws_bytes, int32_t* status, void* stream);

/* The Ros2 step without a workspace argument and without any allocation
 * (no hipMalloc, hipFree, stream or event creation).  No counterpart in the
 * reference: it is for callers that cannot allocate, such as code under HIP
 * graph capture and tight serving loops.  The workspace is a file-scope
 * device array of VK_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends (160 MiB: one column of nz = 200, ni <= 96 in latency mode, or three
 * such columns with the tile factor; vk_ros2_workspace_bytes says what a
 * batch needs).  A batch needing more is refused with VK_ERR_UNSUPPORTED.
 * Same arguments and results as vk_ros2_step_f64, bit for bit; the size
 * buckets of an uneven batch run in order on the given stream.  Every call
 * on a GPU shares its array: calls that overlap on one GPU race. */
#define VK_NOALLOC_SCRATCH_BYTES (160ull << 20)
int vk_ros2_step_noalloc_f64(vk_batch b, const double* y, const double* k, const double* atm,
                             const double* dt, const uint8_t* fix_mask, double* sol,
                             double* delta, int32_t* status, void* stream);

/* Step error of the fixed-dt loop (integrate.py:30-33): out[c * stride] =
 * max over column c of delta / max(|sol|, 1e-30); NaN propagates. */
int vk_step_error_f64(vk_batch b, const double* sol, const double* delta, double* out,
                      int64_t stride, void* stream);

/* f(y) = chemistry + diffusion at y (off_s). */
int vk_residual_f64(vk_batch b, const double* y, const double* k, const double* atm,
                    double* out, void* stream);
/* J_y blocks without c0: diag (natural, off_s*ni), sup_d = B+Bi, sub_d = C+Ci. */
int vk_jacobian_blocks_f64(vk_batch b, const double* y, const double* k, const double* atm,
                           double* diag, double* sup_d, double* sub_d, void* stream);

/* Parameter cotangents of the implicit gradient, the VJP of _ssi_bwd and
 * _ss_bwd through f = chem_rhs + diffusion (steady_state_grad.py:214-233,
 * 306-327), negated as the implicit-function theorem needs:
 *   cot_k   (S, nz) at off_k:     -(d f / d k)^T lam
 *   cot_atm record layout at off_a: -(d f / d theta)^T lam for every
 *           floating leaf of the atmosphere record (Kzz, dzi, vz, Hpi, Ti,
 *           Tco, g, M, ms, alpha, top_flux, bot_flux, bot_vdep, Dzz, vm,
 *           vs); the gas mask and the mode flags get 0.
 * lam at off_s (the adjoint state J_y^{-T} v).  Either output may be NULL. */
int vk_ss_vjp_f64(vk_batch b, const double* y, const double* k, const double* atm,
                  const double* lam, double* cot_k, double* cot_atm, void* stream);

/* ---- launch-shape tuning (no counterpart in the reference) ----
 * Knob VK_TUNE_FACTOR picks the block-factor kernel per SL bucket
 * (SL = ceil(ni/16)): VK_FACTOR_TILE (one 256-thread workgroup per column,
 * 16x16 register tiles, any SL <= 7), VK_FACTOR_ROWS (half a row per lane,
 * 64-192 lanes per column, SL <= 6; bitwise equal to TILE) or
 * VK_FACTOR_BLOCKED (16-column panels eliminated in LDS, the rest updated
 * by rank-16 FP64 MFMA products; same pivots, rounding differs in those
 * products), VK_FACTOR_TILE2 (TILE's algorithm with two waves per column,
 * two rows per lane per tile row; bitwise equal to TILE) or VK_FACTOR_WAVE
 * (the same with one wave per column, four rows per lane per tile row; no
 * barrier inside a column) or VK_FACTOR_TILE8 (eight waves per column, half
 * the rows per lane; even SL) or VK_FACTOR_CR (block cyclic reduction for
 * batches of a few columns: about log2(nz) levels of one workgroup per block
 * row instead of one workgroup walking nz rows; each block inverted with the
 * same partial pivoting, the rows combined in another order, so results
 * agree with TILE to the system's conditioning, not bitwise.  Only
 * vk_ros2_step_f64 and the steady-state attempt use it; it needs the larger
 * workspace vk_ros2_workspace_bytes() reports while it is selected).
 * Process-wide; set it between launches. */
#define VK_TUNE_FACTOR 0
#define VK_FACTOR_TILE 0
#define VK_FACTOR_ROWS 1
#define VK_FACTOR_BLOCKED 2
#define VK_FACTOR_TILE2 3
#define VK_FACTOR_WAVE 4
#define VK_FACTOR_TILE8 5
#define VK_FACTOR_CR 6
/* Knob VK_TUNE_MOLPRE (sl ignored): 1 (default) has vk_ros2_step_f64 compute
 * the molecular-diffusion factors that do not depend on y once per step;
 * 0 forms them inside each chemistry launch.  Results are bitwise equal. */
#define VK_TUNE_MOLPRE 1
/* Knob VK_TUNE_REFINE (sl ignored): one step of iterative refinement of the
 * Ros2 step's block solves.  A solve with explicit block inverses is not
 * backward stable; the refined one is, as the reference's lu_solve is
 * (solver.py:91-112).  The block-Thomas factors refine each block solve of
 * the backward sweep, x_j = X_j t_j; x_j += X_j (t_j - A'_j x_j) with A'_j
 * the Schur complement X_j inverts (the tile factor stores each A'_j beside
 * its inverse, so a refined sweep reads one more block stack; the other
 * variants form A'_j x_j from D_j and X_{j-1}); a column where some layer's
 * correction exceeded
 * that layer's solution (an inverse with no correct digit, cond(A'_j) of
 * 1e12 and more) then takes one whole-column step, residual and correction
 * solve, which the other columns skip.  The cyclic-reduction factor takes
 * the whole-column step for every column.  0: none; 1: the stage-1 solve;
 * 2 (default): both stages.  Unrefined, a step leaves up to ~1e-9 relative
 * error on cells that the reference's LU determines to 1e-12. */
#define VK_TUNE_REFINE 2
/* Knob VK_TUNE_SPLIT (sl ignored): 1..8 column chunks per size bucket in
 * vk_ros2_step_f64 and the steady-state attempt (default 1; a chunk holds at
 * least 128 columns).  Each chunk runs on its own stream, so one chunk's
 * latency-bound block factor can overlap another's bandwidth-bound
 * chemistry and sweeps; measured on MI355X the factor then slows more than
 * the overlap gains.  Results are bitwise independent of the split. */
#define VK_TUNE_SPLIT 3
/* Knob VK_TUNE_KDIRECT (sl ignored): 1 has vk_ros2_step_f64 and the
 * steady-state attempt read the rate tables k in their (S, nz) layout from
 * the chemistry launches; 0 (default) first copies them layer-major (one
 * contiguous row per layer).  Results are bitwise equal. */
#define VK_TUNE_KDIRECT 4
/* Knob VK_TUNE_HOLDX (sl ignored): the refined block solves on the Schur
 * complements the tile factor stores.  1 (default): each layer reads X_j and
 * A'_j once and keeps X_j in registers for both of its products (two block
 * reads per layer, two columns per CU at SL 5-6); 0: X_j is read again for
 * the correction (three block reads, four columns per CU).  Results are
 * bitwise equal. */
#define VK_TUNE_HOLDX 5
int vk_tune_set(int knob, int sl, int value);
int vk_tune_get(int knob, int sl); /* -1 for an unknown knob / bucket */

/* ---- adaptive steady-state runs with the step control on the device ----
 * Replaces the accept/reject/dt logic of outer_loop.body_fn
 * (outer_loop.py:181-250, 741-768, 847-917, 1031-1038) and its stopping
 * rule (cond_fn, outer_loop.py:672-713; relative change of _f_dy,
 * outer_loop.py:2163-2184).  Defaults: vulcan_cfg.py:125-152. */
#define VK_MAX_ATOMS 8

typedef struct {
  double rtol, dt_var_min, dt_var_max, dt_min, dt_max, atol, mtol, pos_cut, nega_cut, loss_eps;
  double trun_min, conv_tol, flux_cri;
  int32_t batch_max_retries, count_min, count_max, na, zero_bot_row, photo_frq;
  int32_t update_frq; /* > 0: flag an atmosphere refresh after accepted steps whose
                         pre-increment count is a multiple (outer_loop.py:813-820) */
  /* conv_mode 0: stop when the last accepted step's largest masked relative
   * change is below conv_tol (_f_dy, outer_loop.py:2163-2184).  1: the
   * reference's stopping rule (_conv_jax and cond_fn, outer_loop.py:636-713,
   * 968-1003): longdy against a ring of the last conv_step accepted states at
   * the entry nearest t * st_factor, longdydt, the strict and loose
   * criteria with slope_min from Kzz and Hp, the stall fallback, runtime and
   * count_max; needs the ring and conv fields of vk_ss_state. */
  int32_t conv_mode, conv_step, conv_stall_window;
  double st_factor, yconv_cri, slope_cri, yconv_min, mtol_conv, runtime;
  /* photolysis cadence switch (outer_loop.py:716-724, 1040-1049): with
   * photo_frq_final > 0 and vk_ss_state.photo_cad, a column starts at
   * photo_frq (the reference's ini_update_photo_frq) and moves to
   * photo_frq_final once longdy < photo_sw_longdy and longdydt <
   * photo_sw_longdydt (the carried values of conv_mode 1; under conv_mode 0
   * the last accepted step's relative change and it over dt) */
  int32_t photo_frq_final;
  double photo_sw_longdy, photo_sw_longdydt;
} vk_ss_params;

/* Device state of a batch of runs.  State-shaped arrays at off_s; compo is
 * the (ni, na) atom composition of each column at off_s; atom arrays are
 * (ncols, VK_MAX_ATOMS); the rest are (ncols).  photo_due and aflux_change
 * may be NULL; with aflux_change, convergence also needs aflux_change <
 * flux_cri (cond_fn, outer_loop.py:688-689). */
typedef struct {
  double *y, *y_prev, *ymix, *sol, *delta_arr;
  const double* compo;
  double *atom_ini, *atom_loss, *atom_loss_prev;
  double *dt, *t, *delta, *rel;
  int32_t *accept, *retry, *nega, *loss, *dcount, *status, *attempts, *photo_due;
  double* aflux_change;
  int32_t* refresh_due; /* (ncols) or NULL: set by the controller (update_frq) */
  /* Per-column constraints on accepted states after the hydrostatic balance
   * (outer_loop.py:855-895), or NULL.  Record at off_pin[c] (device int64):
   * flags (VK_PIN_*), e_idx, charge (ni; 0 at e_idx), bottom densities (ni),
   * bottom mixing ratios of pinned species (ni; NaN where not pinned), then
   * the H2/He pin: h2_idx, he_idx, pin_time, and its state (pinned 0/1,
   * H2 and He bottom mixing ratios), which every attempt updates. */
  double* pin;
  const int64_t* off_pin;
  /* conv_mode 1 only: ring_y (conv_step states of each column at
   * conv_step * off_s), ring_t (ncols, conv_step), conv (ncols, 4: longdy,
   * longdydt, smallest longdy seen, accepted steps since it last dropped 5 %;
   * start at 0, 0, +inf, 0), the layer scale heights Hp (nz) of column c at
   * hp + off_hp[c] (slope_min; may point at a refresh geometry array), and the
   * mask of species the criterion ignores (ni, nonzero = ignored;
   * conver_ignore) at conv_ign + off_ign[c]. */
  double* ring_y;
  double* ring_t;
  double* conv;
  const double* hp;
  const int64_t* off_hp;
  const double* conv_ign;
  const int64_t* off_ign;
  /* (ncols) current photolysis cadence of each column, or NULL (fixed
   * photo_frq); vk_ss_init sets it to photo_frq */
  int32_t* photo_cad;
} vk_ss_state;
#define VK_PIN_CHARGE 1  /* e = -sum_i charge_i y_i in every layer */
#define VK_PIN_ALL_BOT 2 /* bottom layer = the bottom densities */
#define VK_PIN_SP_BOT 4  /* bottom layer of pinned species = mix * n0[0] */
#define VK_PIN_H2HE 8    /* once t > pin_time: bottom H2 and He held at the mixing
                            ratios of that moment (outer_loop.py:871-895) */

/* ymix, y_prev, atom_ini, zeroed counters and status; y and dt must be set. */
int vk_ss_init_f64(vk_batch b, const double* atm, const vk_ss_params* p, const vk_ss_state* s,
                   void* stream);
/* One attempt for every column whose status has no DONE bit: a Ros2 step at
 * (y, dt) into sol/delta_arr, then the device controller.  No host sync, no
 * allocation: safe to capture in a hipGraph and replay. */
/* vk_ss_init_f64 for the columns [first, first + count) only (a refilled slot). */
int vk_ss_init_range_f64(vk_batch b, int first, int count, const double* atm,
                         const vk_ss_params* p, const vk_ss_state* s, void* stream);
int vk_ss_attempt_f64(vk_batch b, const double* k, const double* atm, const vk_ss_params* p,
                      const vk_ss_state* s, void* ws, size_t ws_bytes, void* stream);

/* ---- atmosphere refresh (atm_refresh.py) -----------------------------------
 * Replaces atm_refresh.update_mu_dz_jax (atm_refresh.py:36-98): mean
 * molecular mass, the hydrostatic column integrated up and down from the
 * reference layer with g = gs (Rp / (Rp + z))^2, and the interface averages;
 * and update_phi_esc_jax (atm_refresh.py:101-122): diffusion-limited escape
 * at the top floored at -max_flux.  Each column has a refresh record
 * (layout from vk_refresh_record_layout: pico (nz+1), mol_mass (ni), escape
 * mask (ni, 1.0 / 0.0), then pref_indx, zco_pref, gs, Rp, max_flux, kb,
 * Navo) at off_r[c], and geometry outputs at off_g[c]: mu, g, Hp, dz (nz
 * each) and zco (nz+1).  The atmosphere record's g, dzi, Hpi and the
 * escaping species' top_flux are rewritten in place (the next step reads
 * them, outer_loop.py:215-216).
 * mode bit 1: geometry from the mixing ratios; bit 2: escape fluxes (from
 * the geometry's top g and Hp).  With sol non-NULL (the in-loop form) the
 * mixing ratios and the top-layer densities are those of the controller's
 * clip of the step output sol against ymix (outer_loop.py:196-224, p's cuts),
 * as the reference refreshes from the clipped state; otherwise ymix and y
 * are used as given.  With r->due non-NULL only flagged columns refresh and
 * their flags are cleared. */
typedef struct {
  const double* rec;
  const int64_t* off_r; /* (ncols) device */
  double* geo;
  const int64_t* off_g; /* (ncols) device */
  int32_t* due;         /* (ncols) or NULL */
} vk_refresh_state;
int64_t vk_refresh_record_layout(int nz, int ni, int64_t* off /* 4 */);
int vk_atm_refresh_f64(vk_batch b, int mode, const double* ymix, const double* y,
                       const double* sol, const vk_ss_params* p, double* atm,
                       const vk_refresh_state* r, void* stream);

/* ---- photolysis (photo.py) -----------------------------------------------
 * One cross-section / stellar-flux setup shared by every column of a call.
 * Per-column arrays: dz (nz) at off_l[c]; tau, w0, sflux, dflux_d, dflux_u,
 * aflux (nz+1, nbin) at off_p[c]; J (n_br, nz) at off_l[c]*n_br; k rows at
 * off_k (S, nz).  Replaces compute_tau_jax (photo.py:84-123),
 * compute_flux_jax (photo.py:126-245; ag0 = 0 and ag0 != 0 branches),
 * compute_J_jax_flat (photo.py:309-351, 377-392) and update_k_with_J
 * (photo.py:455-478).  The J contraction (n_br x nbin) . (nbin x nz) runs on
 * FP64 MFMA.  Temperature-dependent cross sections (photo.py:107-111,
 * 156-160, 334-347) are per column, at off_l[c] * n * nbin: absp_T_cross
 * (n_absT, nz, nbin) for the absorbers absp_T_idx, cross_J_T (n_brT, nz,
 * nbin) for the branches whose J_T (n_brT, nz, at off_l[c] * n_brT in the
 * state) fills the k rows branch_T_re_idx where branch_T_active. */
typedef struct {
  int32_t nbin, n_abs, n_scat, n_br, din12, ag0_zero;
  double dbin1, dbin2, mu_zenith, edd, ag0, hc, f_diurnal, flux_atol;
  const double *bins, *sflux_top, *absp_cross, *scat_cross, *cross_J, *dz;
  const int32_t *absp_idx, *scat_idx, *branch_re_idx, *branch_active;
  const int64_t *off_p, *off_l;
  int32_t n_absT, n_brT;
  const double *absp_T_cross, *cross_J_T;
  const int32_t *absp_T_idx, *branch_T_re_idx, *branch_T_active;
} vk_photo_setup;

/* due (may be NULL): per-column flag, columns with 0 are skipped; status
 * (may be NULL): columns with a DONE bit are skipped; aflux_change (may be
 * NULL): per-column max relative aflux change (outer_loop.py:330-336). */
typedef struct {
  double *tau, *w0, *sflux, *dflux_d, *dflux_u, *aflux, *J, *aflux_change;
  const int32_t *due, *status;
  double* J_T; /* (n_brT, nz) per column; may be NULL when n_brT == 0 */
} vk_photo_state;

/* parts: 1 = optics (tau, w0 from y), 2 = flux sweeps, 4 = J (+ k rows when
 * k != NULL); OR them for the in-loop update. */
int vk_photo_f64(vk_batch b, int parts, const vk_photo_setup* setup, const vk_photo_state* state,
                 const double* y, double* k, void* stream);

/* ---- rate construction k(T, M) ----
 * Restates the reference's host-side rate build on the device:
 * rates.compute_forward_k (rates.py:61-122), rates.apply_lowT_caps
 * (rates.py:142-167), gibbs.gibbs_sp_vector (gibbs.py:41-74),
 * gibbs.K_eq_array (gibbs.py:77-128), gibbs.fill_reverse_k (gibbs.py:131-154)
 * and rates.build_rate_array (rates.py:185-208).
 *
 * vk_rates_create takes the network's rows classified once on the host:
 *   form[i]    VK_KF_* forward form of row i (VK_KF_ZERO for reverse rows,
 *              photo/ion/condensation/radiative rows and rows the reference
 *              leaves at zero);
 *   lowt[i]    VK_LOWT_* low-T cap of forward row i;
 *   has_rev[i] forward row i has a thermal reverse at i+1 (K_eq computed);
 *   par        (nr+1, 6): a, n, E, a_inf, n_inf, E_inf;
 *   tables     (nr+1, T) reactant/product species (PAD = ni) and stoich,
 *              T <= 4, stoich exact in fp32, ni <= 255;
 *   nasa9      (ni, 2, 10) NASA-9 coefficients, low-T range first (NULL:
 *              no Gibbs / K_eq / build);
 *   zero_rows  (nr+1) rows the build and the reverse fill zero (the remove
 *              list, paired or not as the caller resolved it), or NULL.
 * vk_rates_eval_f64 evaluates one column on nz layers, out (rows, nz):
 *   VK_RATES_FORWARD  compute_forward_k (+ caps when lowt_on);
 *   VK_RATES_CAPS     apply_lowT_caps of in_a (rows, nz);
 *   VK_RATES_GIBBS    gibbs_sp_vector, out (ni, nz);
 *   VK_RATES_KEQ      K_eq_array from g/RT in in_a (ni, nz);
 *   VK_RATES_FILL     fill_reverse_k of k in_a and K_eq in_b;
 *   VK_RATES_BUILD    forward (+ caps), reverse fill from NASA-9, zero rows.
 * vk_batch_set_rates gives a plain batch one rates handle per network (in
 * the batch's network order); vk_rates_build_batch_f64 then writes every
 * column's k (S, nz) at off_k from the T and M of its atmosphere record. */
typedef struct vk_rates_s* vk_rates;
#define VK_KF_ZERO 0
#define VK_KF_ARRHENIUS 1
#define VK_KF_LINDEMANN 2
#define VK_KF_TROE_OH_CH3 3
#define VK_LOWT_NONE 0
#define VK_LOWT_CH3 1  /* H + CH3 + M -> CH4 + M below 277.5 K */
#define VK_LOWT_C2H4 2 /* H + C2H4 + M -> C2H5 + M below 300 K */
#define VK_LOWT_C2H5 3 /* H + C2H5 + M -> C2H6 + M below 200 K */
#define VK_RATES_FORWARD 0
#define VK_RATES_CAPS 1
#define VK_RATES_GIBBS 2
#define VK_RATES_KEQ 3
#define VK_RATES_FILL 4
#define VK_RATES_BUILD 5
int vk_rates_create(int nr, int ni, int T, int stop_rev_indx, const int32_t* form,
                    const int32_t* lowt, const int32_t* has_rev, const double* par,
                    const int64_t* reactant_idx, const double* reactant_stoich,
                    const int64_t* product_idx, const double* product_stoich,
                    const double* nasa9, const uint8_t* zero_rows, vk_rates* out);
int vk_rates_destroy(vk_rates r);
int vk_rates_eval_f64(vk_rates r, int what, int nz, const double* T, const double* M,
                      int lowt_on, const double* in_a, const double* in_b, double* out,
                      void* stream);
/* vk_rates_vjp_f64: the cotangents of the full build's inputs,
 * T_bar[z] = sum_i k_bar[i, z] dk[i, z]/dT[z] and M_bar likewise, for
 * k = VK_RATES_BUILD(T, M).  No counterpart in the reference, whose NumPy
 * rate build sits outside its autodiff; it is for callers that need
 * gradients through k(T) and M(T) (the steady-state sensitivity to T(z)).
 * k_bar (nr+1, nz); either output may be NULL; the sums over rows are in a
 * fixed order. */
int vk_rates_vjp_f64(vk_rates r, int nz, const double* T, const double* M, int lowt_on,
                     const double* k_bar, double* T_bar, double* M_bar, void* stream);
int vk_batch_set_rates(vk_batch b, const vk_rates* per_net, int n);
int vk_rates_build_batch_f64(vk_batch b, int lowt_on, const double* atm, double* k, void* stream);

/* Kernel timing with HIP events recorded on the launch stream around each
 * kernel of vk_ros2_step_f64 (no counterpart in the reference, which has no
 * device timing).  Kernel ids: 0 chemistry+LHS, 1 block factor, 2 stage-1
 * back sweep, 3 stage-2 rhs, 4 stage-2 solve, 5 step control, 6 per-step
 * preparation (rate-table transpose and the molecular-diffusion factors that
 * do not depend on y).  vk_timing_read synchronises
 * on the recorded events and returns accumulated milliseconds and launch
 * counts per id.  Not thread-safe; for benchmarks. */
int vk_timing_enable(int on);
int vk_timing_reset(void);
int vk_timing_read(double* ms, long long* count, int n);

#ifdef __cplusplus
}
#endif

#endif /* VULCAN_ROS2_H */
