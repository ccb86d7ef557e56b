#!/usr/bin/env python
"""Ros2 column-step throughput on one MI355X (FP64, nz=150, ni=93).

One bench "step" is one full Ros2 step (chemistry + Jacobian + block-Thomas
factor + two block solves + combine) for every column of a batch of
independent columns of the nz=150 / ni=93 / 2461-rate-row workload, inputs
resident in HBM.  Prints one JSON line; see the DESIGN notes for the
roofline and CPU-baseline definitions.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

FACTOR_CR = 6    # vk_tune_set(VK_TUNE_FACTOR, sl, VK_FACTOR_CR): cyclic-reduction factor
VK_TUNE_REFINE, VK_TUNE_SPLIT, VK_TUNE_KDIRECT, VK_TUNE_HOLDX = 2, 3, 4, 5
# The synthetic initial states put trace species at 1e-20 of the layer: at the
# reference's default dt_min (1e-14) the error estimate exceeds rtol and the
# force-accept rule (which keeps y_prev) freezes a column; the steady-state
# legs run with dt_min = 1e-30 (a run-configuration value, vulcan_cfg.dt_min).
SERVE_DT_MIN = 1e-30

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "jax-vulcan_amd")]

NZ, NI, NREV, NPHOTO = 150, 93, 1200, 60
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector = matrix)
PEAK_HBM_GBS = 8000.0


def factor_flops(nz: int, n: int) -> float:
    """Algorithmic FLOPs of one block-Thomas factor of one column:
    nz LU's (2/3 n^3) + (nz-1) inverses (4/3 n^3) + Schur updates (3 n^2)."""
    return nz * (2.0 / 3.0) * n ** 3 + (nz - 1) * ((4.0 / 3.0) * n ** 3 + 3.0 * n ** 2)


def step_flops(nz: int, n: int, S: int) -> float:
    """Algorithmic FLOPs of one Ros2 column-step (factor + 2 solves + chemistry)."""
    f_sol = 2.0 * ((nz - 1) * (2 * n * n + 2 * n) * 2 + 2 * n * n)
    return factor_flops(nz, n) + f_sol + 44.0 * nz * S


def factor_traffic(B: int):
    """HBM bytes per factor launch from the committed rocprofv3 counter runs
    (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a 1 GiB copy) at
    1024 columns, scaled to B columns; None when the summary is absent."""
    f = ROOT / "profiles" / "ros2_step_nz150_ni93_b1024_hbm_counters.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("factor_traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1]) * B / 1024)
    return None


def bench_rates(B: int, net, atm, reps: int = 20, cpu_cols: int = 20, no_cpu: bool = False):
    """Batched rate construction k(T, M) (forward rates, NASA-9 K_eq, reverse
    fill) for B columns of the step's network, T and M from each column's
    packed atmosphere record, timed with HIP events on the launch stream;
    the NumPy oracle times a few columns on the host."""
    import numpy as np
    import torch
    from vulcan_amd import core, rates, synth
    rng = np.random.default_rng(77)
    rnet = synth.random_rate_network(rng, net, NREV, NPHOTO)
    c9 = synth.random_nasa9(rng, NI)
    b = core.uniform_batch(B, NZ, net)
    out = torch.empty(b.k_len, dtype=torch.float64, device="cuda")
    rates.build_rate_batch(b, [rnet], [c9], atm, False, None, out=out)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        rates.build_rate_batch(b, [rnet], [c9], atm, False, None, out=out)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = b.k_len * 8
    gbs = nbytes / (ms * 1e-3) / 1e9
    res = {"workload": f"build_rate_array for {B} columns, nz={NZ}, ni={NI}, "
                       f"{net.nr + 1} rate rows (random Arrhenius/Lindemann/Troe/cap rows, "
                       f"random NASA-9 fits)",
           "ms_per_batch": round(ms, 4), "column_builds_per_s": round(B / (ms * 1e-3), 1),
           "bytes_written": nbytes, "write_GBps": round(gbs, 1),
           "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)}
    if not no_cpu:
        from oracle import rates_np as O
        rec_len = atm.numel() // B
        _, off = core.atm_record_layout(NZ, NI)
        r0 = atm[:rec_len].cpu().numpy()
        T = r0[off[5]:off[5] + NZ]
        M = r0[off[7]:off[7] + NZ]
        O.build_rate_array(rnet, T, M, c9)
        t0 = time.perf_counter()
        with np.errstate(all="ignore"):
            for _ in range(cpu_cols):
                O.build_rate_array(rnet, T, M, c9)
        cpu_s = (time.perf_counter() - t0) / cpu_cols
        res["cpu_oracle_column_builds_per_s"] = round(1.0 / cpu_s, 2)
        res["cpu_sample"] = f"{cpu_cols} columns, NumPy oracle, 1 thread"
    return res


def bench_config4(reps: int = 5):
    """The differentiation workload: the steady state of the open nz=150,
    ni=93, 1200-reaction column (synth.config4) by pseudo-transient
    continuation on the cyclic-reduction factor, then the implicit gradient
    of L = sum log ymix of 5 species with respect to every leaf (transposed
    block Thomas + parameter VJP), timed on both factors."""
    import numpy as np
    import torch
    from vulcan_amd import steady_state_grad as G
    from vulcan_amd import synth
    pb = synth.config4()
    k0, atm0 = pb.leaves(pb.T0)
    spt = torch.as_tensor(pb.loss_sp, device="cuda")

    def loss(y):
        ymix = y / y.sum(dim=1, keepdim=True)
        return ymix[:, spt].log().sum()

    out = {"workload": "open column nz=150, ni=93, 1200 reactions (2401 rate rows), bottom "
                       "fluxes of 10 species; L = sum over layers of log ymix of 5 species; "
                       "gradient wrt every k_arr entry and atmosphere leaf (1200 log k, 150 T(z) "
                       "and 10 bottom fluxes by the chain rule)"}
    try:
        G.set_latency_mode(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ys, it, res = G.pseudo_transient_steady_state(pb.initial_state(), k0, atm0, pb.net)
        torch.cuda.synchronize()
        out["forward_solve_s"] = round(time.perf_counter() - t0, 3)
        out["forward_iterations"] = it
        out["residual_inf"] = res
        inputs = G.build_steady_state_inputs(k0, atm0)
        for mode in (False, True):
            G.set_latency_mode(mode)
            G.steady_state_value_and_grad(loss, inputs, ys, pb.net, residual_rtol=1e-9)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                _, gi = G.steady_state_value_and_grad(loss, inputs, ys, pb.net,
                                                      residual_rtol=1e-9)
            torch.cuda.synchronize()
            key = "gradient_ms_cyclic_reduction" if mode else "gradient_ms_block_thomas"
            out[key] = round((time.perf_counter() - t0) / reps * 1e3, 2)
        gk = gi.k_arr.detach().cpu().numpy()
        g_m = (k0[1::2] * gk[1::2] + k0[2::2] * gk[2::2]).sum(axis=1)
        out["max_abs_dL_dlogk"] = float(np.abs(g_m).max())
    finally:
        G.set_latency_mode(False)
    return out


def bench_config2(photo_batch: int):
    """One full steady-state solve of the nz=150, ni=93, 2461-row column
    (independent K_eq) with two-stream photolysis over 2500 bins refreshed in
    the loop every 5 accepted steps, on the cyclic-reduction factor; and the
    photolysis update (tau, two-stream, J on MFMA, k rows) of photo_batch
    such columns in one launch set."""
    import ctypes as C
    import numpy as np
    import torch
    from vulcan_amd import _lib, core, photo, synth
    from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
    c = synth.config2()
    r = SteadyStateRunner([c], RunConfig(dt_min=SERVE_DT_MIN), latency_mode=True)
    r.run(max_attempts=40, check_every=20, use_graph=False)     # warm the code paths
    r.reset_to_initial()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r.run(max_attempts=100000, check_every=50, use_graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    res = r.results()
    out = {"workload": "nz=150, ni=93, 1200 reversible (independent K_eq) + 60 photolysis rows, "
                       "nbin=2500 two-stream refreshed every 5 accepted steps, adaptive Ros2 to "
                       "relative change < 1e-4 or 5000 accepted steps, dt_min 1e-30",
           "factor_kernel": "cr", "wall_s": round(wall, 3), "attempts": int(res["attempts"][0]),
           "accepted": int(res["accept"][0]),
           "converged": bool(res["status"][0] & 4), "capped": bool(res["status"][0] & 8),
           "ms_per_attempt": round(wall / max(int(res["attempts"][0]), 1) * 1e3, 3),
           "t_end_s": float(res["t"][0]), "last_rel_change": float(res["rel"][0])}
    # batched photolysis update
    B = photo_batch
    b = core.uniform_batch(B, c.nz, c.net)
    ph = photo._ph_dict(c.photo)
    setup = photo._Setup(ph, [c.nz] * B, [c.dz] * B, branch_re_idx=ph["branch_re_idx"],
                         branch_active=ph["branch_active"])
    st = photo._State(setup.plen, setup.llen, setup.n_br, B)
    cst = st.c()
    y = torch.as_tensor(np.tile(c.y, (B, 1)), dtype=torch.float64, device="cuda").reshape(-1)
    k = torch.as_tensor(np.tile(c.k, (B, 1)), dtype=torch.float64, device="cuda").reshape(-1)
    lib = _lib.load()
    stream = torch.cuda.current_stream().cuda_stream

    def upd():
        _lib.check(lib.vk_photo_f64(b.handle, 7, C.byref(setup.c), C.byref(cst), y.data_ptr(),
                                    k.data_ptr(), stream), "vk_photo_f64")
    upd()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        upd()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbin, n_br, nz = setup.nbin, setup.n_br, c.nz
    out["photo_update_batch"] = {
        "columns": B, "ms_per_update": round(ms, 4),
        "column_updates_per_s": round(B / (ms * 1e-3), 1),
        "J_contraction_flop": 2 * n_br * nz * nbin * B,
        "flux_state_bytes": 7 * (nz + 1) * nbin * 8 * B}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--single-steps", type=int, default=5)
    ap.add_argument("--cpu-sample", type=int, default=64,
                    help="column-steps timed on the host for the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--uneven", type=int, default=512,
                    help="columns of the uneven-size batch leg (nz 60..200, ni 50/69/93; "
                         "0 skips it)")
    ap.add_argument("--uneven-attempts", type=int, default=10,
                    help="steady-state attempts per column timed in the uneven leg")
    ap.add_argument("--serve", type=int, default=64,
                    help="serving leg: this many steady-state solve requests (sizes of the "
                         "uneven batch, Poisson arrivals at --serve-rate per s, 200 GB budget)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-factor", choices=("cr", "auto"), default="cr",
                    help="block factor of the serving pools: cyclic reduction (latency mode) "
                         "or the throughput kernels (--factor)")
    ap.add_argument("--serve-alone", type=int, default=8,
                    help="requests also solved one at a time, for the comparison")
    ap.add_argument("--serve-budget-gb", type=float, default=0.0,
                    help="HBM budget of the serving pools (0: 200 GB scaled by requests / "
                         "10000, the full workload's budget per request)")
    ap.add_argument("--config2", action=argparse.BooleanOptionalAction, default=True,
                    help="time one full steady-state solve of the nz=150, ni=93 column with "
                         "in-loop photolysis (nbin=2500), and a batched photolysis update")
    ap.add_argument("--metrics", default=None, metavar="PATH",
                    help="append JSON Lines records (legs, serving requests, result) to PATH")
    ap.add_argument("--config4", action=argparse.BooleanOptionalAction, default=True,
                    help="time the differentiation workload (steady state + implicit gradient)")
    ap.add_argument("--photo-batch", type=int, default=256,
                    help="columns of the batched photolysis-update timing")
    ap.add_argument("--factor", choices=("auto", "tile", "rows", "blocked", "tile2", "wave", "tile8", "cr"),
                    default="auto",
                    help="block-factor kernel for the nz=150 / ni=93 bucket (auto: the "
                         "library default)")
    ap.add_argument("--split", type=int, default=0,
                    help="column chunks per size bucket, each on its own stream "
                         "(VK_TUNE_SPLIT; 0: the library default)")
    ap.add_argument("--kdirect", type=int, default=-1,
                    help="chemistry reads k in its (S, nz) layout (VK_TUNE_KDIRECT 1) or a "
                         "layer-major copy (0); -1 keeps the library default")
    ap.add_argument("--holdx", type=int, default=-1,
                    help="refined sweeps keep X_j in registers for both of its products "
                         "(VK_TUNE_HOLDX 1) or read it again (0); -1 keeps the library default")
    ap.add_argument("--refine", type=int, default=-1,
                    help="refined block solves (VK_TUNE_REFINE: 0 none, 1 stage 1, 2 both; "
                         "-1: the library default)")
    ap.add_argument("--conv-cost", type=int, default=0,
                    help="columns (nz=150, ni=93) for timing one attempt under the "
                         "relative-change rule and under the reference's longdy rule "
                         "(ring of 500); 0 skips it")
    ap.add_argument("--solve-latency", action="store_true",
                    help="time one transposed factor + solve of a random nz=150, ni=93 "
                         "block system (the adjoint solve of one column), block Thomas "
                         "against cyclic reduction")
    ap.add_argument("--no-rates", action="store_true",
                    help="skip the rate-construction measurement")
    ap.add_argument("--rates-only", action="store_true",
                    help="only measure the batched rate construction k(T, M)")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run a plain elementwise copy of GIB GiB (reads GIB, writes "
                         "GIB): the reference kernel that scales the HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(columns are independent: multi-GPU runs shard columns with no exchange)",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from vulcan_amd import _lib, core, synth
    from vulcan_amd.metrics import open_log
    mlog = open_log(args.metrics)

    if args.copy_probe > 0.0:
        n = int(args.copy_probe * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n}))
        return 0

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    lib = _lib.load()
    if args.factor != "auto":
        _lib.check(lib.vk_tune_set(0, (NI + 15) // 16,
                                   {"tile": 0, "rows": 1, "blocked": 2, "tile2": 3, "wave": 4, "tile8": 5,
                                    "cr": FACTOR_CR}[args.factor]),
                   "vk_tune_set")
    if args.split > 0:
        _lib.check(lib.vk_tune_set(VK_TUNE_SPLIT, 0, args.split), "vk_tune_set")
    if args.refine >= 0:
        _lib.check(lib.vk_tune_set(VK_TUNE_REFINE, 0, args.refine), "vk_tune_set")
    if args.kdirect >= 0:
        _lib.check(lib.vk_tune_set(VK_TUNE_KDIRECT, 0, args.kdirect), "vk_tune_set")
    if args.holdx >= 0:
        _lib.check(lib.vk_tune_set(VK_TUNE_HOLDX, 0, args.holdx), "vk_tune_set")
    B = args.batch
    t_gen = time.time()
    cols = synth.columns_same_network(1002, B, NZ, NI, NREV, NPHOTO, dt=1e-10, name="bench")
    net = cols[0].net
    S = net.nr + 1
    y = torch.tensor(np.stack([c.y for c in cols]), dtype=torch.float64, device=dev)
    k = torch.tensor(np.stack([c.k for c in cols]), dtype=torch.float64, device=dev)
    atm = torch.cat([core.pack_atm_record(c.atm) for c in cols]).to(dev)
    dt = torch.full((B,), 1e-10, dtype=torch.float64, device=dev)
    print(f"[bench] generated {B} columns in {time.time() - t_gen:.1f}s", file=sys.stderr, flush=True)

    if args.rates_only:
        rb = bench_rates(B, net, atm, reps=max(args.steps, 1), no_cpu=args.no_cpu)
        print(json.dumps({"metric": "rate-construction column-builds/s",
                          "value": rb["column_builds_per_s"], "unit": "column-builds/s",
                          "n_gpus": 1, "dtype": "f64", "data": "synthetic", "rate_build": rb}))
        return 0

    def make(Bn):
        b = core.uniform_batch(Bn, NZ, net)
        ws = b.workspace()
        sol = torch.empty((Bn, NZ, NI), dtype=torch.float64, device=dev)
        delta = torch.empty_like(sol)
        status = torch.zeros(Bn, dtype=torch.int32, device=dev)
        return b, ws, sol, delta, status

    def stepper(Bn):
        b, ws, sol, delta, status = make(Bn)
        st = torch.cuda.current_stream().cuda_stream

        def run():
            rc = lib.vk_ros2_step_f64(b.handle, y.data_ptr(), k.data_ptr(), atm.data_ptr(),
                                      dt.data_ptr(), None, sol.data_ptr(), delta.data_ptr(),
                                      ws.data_ptr(), ws.numel(), status.data_ptr(), st)
            _lib.check(rc, "vk_ros2_step_f64")
        return run, sol, status

    # ---- batched throughput ------------------------------------------------
    main_factor = args.factor
    if args.factor == "auto":
        from vulcan_amd import autotune
        main_factor = autotune.tune_columns(cols, verbose=True)[(NI + 15) // 16]
    run, sol, status = stepper(B)
    for _ in range(args.warmup):
        run()
    torch.cuda.synchronize()
    lib.vk_timing_reset()
    lib.vk_timing_enable(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    lib.vk_timing_enable(0)
    ms = (C.c_double * 7)()
    cnt = (C.c_longlong * 7)()
    lib.vk_timing_read(ms, cnt, 7)
    kern_ms = [ms[i] / max(cnt[i], 1) for i in range(7)]
    elapsed = t1 - t0
    value = B * args.steps / elapsed
    finite = bool(torch.isfinite(sol).all().item())
    n_singular = int((status != 0).sum().item())
    print(f"[bench] batch {B}: {elapsed / args.steps * 1e3:.2f} ms/step, "
          f"{value:.1f} column-steps/s; kernels ms {['%.3f' % v for v in kern_ms[:5]]} "
          f"prep {kern_ms[6]:.3f}",
          file=sys.stderr, flush=True)

    # ---- plain single-column kernel (the tile factor) --------------------------
    sl_main = (NI + 15) // 16
    prev = lib.vk_tune_get(0, sl_main)
    _lib.check(lib.vk_tune_set(0, sl_main, 0), "vk_tune_set")
    try:
        run1, _, _ = stepper(1)
        run1()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.single_steps):
            run1()
        torch.cuda.synchronize()
        single = args.single_steps / (time.perf_counter() - t0)
    finally:
        lib.vk_tune_set(0, sl_main, prev)
    print(f"[bench] single column: {single:.1f} column-steps/s", file=sys.stderr, flush=True)
    # the same column with the cyclic-reduction factor (latency mode)
    _lib.check(lib.vk_tune_set(0, sl_main, FACTOR_CR), "vk_tune_set")
    try:
        run1c, _, _ = stepper(1)
        run1c()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.single_steps):
            run1c()
        torch.cuda.synchronize()
        single_cr = args.single_steps / (time.perf_counter() - t0)
    finally:
        lib.vk_tune_set(0, sl_main, prev)
    print(f"[bench] single column, cyclic-reduction factor: {single_cr:.1f} column-steps/s",
          file=sys.stderr, flush=True)

    # ---- uneven batch: adaptive steady-state attempts, one graph replay --------
    uneven = None
    if args.uneven > 0:
        from vulcan_amd import autotune
        from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
        t_gen = time.time()
        cols3 = synth.config3(1003, args.uneven)
        print(f"[bench] uneven batch: {len(cols3)} columns generated in {time.time() - t_gen:.1f}s",
              file=sys.stderr, flush=True)
        chosen = autotune.tune_columns(cols3, verbose=True) if args.factor == "auto" else {}
        runner = SteadyStateRunner(cols3, RunConfig())
        runner.capture(args.uneven_attempts)
        runner.graph[0].replay()  # warm replay
        torch.cuda.synchronize()
        runner.reset_to_initial()
        torch.cuda.synchronize()
        a0 = runner.attempts.sum().item()
        t0 = time.perf_counter()
        runner.graph[0].replay()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        col_steps = runner.attempts.sum().item() - a0
        uneven = {"columns": len(cols3), "attempts_per_column": args.uneven_attempts,
                  "column_steps": int(col_steps), "seconds": round(el, 4),
                  "column_steps_per_s": round(col_steps / el, 2),
                  "ms_per_attempt": round(el / args.uneven_attempts * 1e3, 3),
                  "sizes": "nz~U{60..200}, ni in {50,69,93}, nr=13*ni, independent columns",
                  "launch": f"one HIP graph of {args.uneven_attempts} attempts",
                  "factor_kernel_by_bucket": {str(k): v for k, v in chosen.items()}}
        print(f"[bench] uneven batch: {col_steps / el:.1f} column-steps/s, "
              f"{el / args.uneven_attempts * 1e3:.2f} ms per attempt", file=sys.stderr, flush=True)
        del runner

    # ---- serving: continuous batching of steady-state solves -----------------
    serving = None
    if args.serve > 0:
        from vulcan_amd import autotune
        from vulcan_amd.outer_loop import RunConfig, SteadyStateRunner
        from vulcan_amd.serving import ServingEngine, latency_summary
        sizes = synth.config3_sizes(1005, args.serve)
        reqs = [synth.make_column(1005 * 100003 + i, nz=a, ni=b_, n_rev=r, dt=1e-10,
                                  name=f"req{i}", consistent=True)
                for i, (a, b_, r) in enumerate(sizes)]
        arrivals = synth.poisson_arrivals(1005, args.serve, args.serve_rate)
        arrivals = arrivals - arrivals[0]
        latency = args.serve_factor == "cr"
        if not latency and args.factor == "auto":
            autotune.tune_columns(reqs, verbose=True)
        budget_gb = args.serve_budget_gb or 200.0 * len(reqs) / 10000.0
        scfg = RunConfig(dt_min=SERVE_DT_MIN)
        eng = ServingEngine(hbm_budget_bytes=budget_gb * 1e9, cfg=scfg, chunk=25,
                            latency_mode=latency)
        t0 = time.perf_counter()
        res = eng.run(reqs, arrivals, report_every=30.0, metrics=mlog)
        wall = time.perf_counter() - t0
        summ = latency_summary(res)
        waits = np.array([r.start - r.arrival for r in res])
        alone = []
        for i in range(min(args.serve_alone, len(reqs))):
            r1 = SteadyStateRunner([reqs[i]], scfg, latency_mode=latency)
            t1 = time.perf_counter()
            r1.run(max_attempts=100000, check_every=25, use_graph=True, report_every=30.0)
            alone.append(time.perf_counter() - t1)
        alone_rate = len(alone) / sum(alone) if alone else None
        serving = {"factor_kernel": args.serve_factor, "requests": len(reqs),
                   "arrival_rate_per_s": args.serve_rate, "hbm_budget_gb": round(budget_gb, 2),
                   "slots": {str(k): v for k, v in eng.slot_counts.items()},
                   "wall_s": round(wall, 3),
                   "p50_latency_s": round(summ["p50_s"], 3), "p99_latency_s": round(summ["p99_s"], 3),
                   "admission_wait_s": {"p50": round(float(np.percentile(waits, 50)), 3),
                                        "p99": round(float(np.percentile(waits, 99)), 3),
                                        "max": round(float(waits.max()), 3)},
                   "throughput_solves_per_s": round(summ["throughput_solves_per_s"], 4),
                   "converged": summ["converged"], "mean_attempts": round(summ["attempts_mean"], 1),
                   "latency_bound": ("time to steady state" if summ["converged"] == len(res)
                                     else f"{len(res) - summ['converged']} requests ran to the "
                                          f"5000-step cap"),
                   "one_at_a_time_solve_s": [round(a, 3) for a in alone],
                   "one_at_a_time_solves_per_s": None if alone_rate is None else round(alone_rate, 4),
                   "served_over_one_at_a_time": (None if not alone_rate else
                                                 round(summ["throughput_solves_per_s"] / alone_rate, 2)),
                   "sizes": "nz~U{60..200}, ni in {50,69,93}, nr=13*ni, consistent K_eq, dt_min "
                            f"{SERVE_DT_MIN:g}; steady state: relative change < 1e-4 (t > 100 s, "
                            "120 accepted steps) or 5000 accepted steps; budget 200 GB x "
                            "requests / 10000"}
        print(f"[bench] serving: {serving}", file=sys.stderr, flush=True)

    # ---- plain steady state with in-loop photolysis (one column) -------------
    config2 = None
    if args.config2:
        config2 = bench_config2(args.photo_batch)
        print(f"[bench] config 2: {config2}", file=sys.stderr, flush=True)

    config4 = None
    if args.config4:
        config4 = bench_config4()
        print(f"[bench] config 4: {config4}", file=sys.stderr, flush=True)

    conv_cost = None
    if args.conv_cost > 0:
        from vulcan_amd.outer_loop import ColumnConv, RunConfig, SteadyStateRunner
        ccols = synth.columns_same_network(1002, args.conv_cost, NZ, NI, 1200)
        cconv = [ColumnConv(Hp=synth.compute_geometry(c.P, c.atm.Tco)["Hp"]) for c in ccols]
        conv_cost = {"columns": args.conv_cost, "conv_step": 500}
        for mode in ("relmax", "longdy"):
            rr = SteadyStateRunner(ccols, RunConfig(conv_mode=mode, conv_step=500),
                                   conv=cconv if mode == "longdy" else None)
            rr.capture(10)
            rr.graph[0].replay()
            torch.cuda.synchronize()
            rr.reset_to_initial()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rr.graph[0].replay()
            torch.cuda.synchronize()
            conv_cost[f"ms_per_attempt_{mode}"] = round((time.perf_counter() - t0) / 10 * 1e3, 3)
            del rr
        print(f"[bench] stopping-rule cost: {conv_cost}", file=sys.stderr, flush=True)

    solve_latency = None
    if args.solve_latency:
        from vulcan_amd import steady_state_grad as G
        rng = np.random.default_rng(1)
        tt = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
        sdiag = tt(rng.normal(size=(NZ, NI, NI)) + 4.0 * np.sqrt(NI) * np.eye(NI)[None])
        ssup, ssub, srhs = (tt(rng.normal(size=(NZ, NI))) for _ in range(3))
        solve_latency = {"system": f"random block tridiagonal, nz={NZ}, ni={NI}, transposed"}
        try:
            for mode in (False, True):
                G.set_latency_mode(mode)
                G._factor_solve(sdiag, ssup, ssub, srhs, net, NZ, True)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(10):
                    G._factor_solve(sdiag, ssup, ssub, srhs, net, NZ, True)
                torch.cuda.synchronize()
                key = "ms_cyclic_reduction" if mode else "ms_block_thomas"
                solve_latency[key] = round((time.perf_counter() - t0) / 10 * 1e3, 3)
        finally:
            G.set_latency_mode(False)
        print(f"[bench] solve latency: {solve_latency}", file=sys.stderr, flush=True)

    rate_build = None
    if not args.no_rates:
        rate_build = bench_rates(B, net, atm, no_cpu=args.no_cpu)
        print(f"[bench] rate build: {rate_build}", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel (block factor) ---------------------
    fac_ms = kern_ms[1]
    fac_flops = B * factor_flops(NZ, NI)
    achieved = fac_flops / (fac_ms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                "traffic": factor_traffic(B),
                "pipe": "FP64 VALU FMA (the tile block factor issues vector FMAs; MI355X's "
                        "FP64 vector and matrix peaks are both 78.6 TFLOP/s)"}

    # ---- CPU baseline: C restatement of the reference step -------------------
    cpu = None
    if not args.no_cpu:
        from oracle import ros2_c
        ncpu = os.cpu_count() or 1
        threads = max(1, min(ncpu, 16))
        n = min(args.cpu_sample, B)
        ys = y[:n].cpu().numpy()
        ks = k[:n].cpu().numpy()
        recs = atm.view(B, -1)[:n].cpu().numpy()
        ros2_c.ros2_steps(ys[:1], ks[:1], recs[:1], 1e-10, net, nthreads=1)  # warm
        t0 = time.perf_counter()
        ros2_c.ros2_steps(ys, ks, recs, 1e-10, net, nthreads=threads)
        cpu_rate = n / (time.perf_counter() - t0)
        m1 = max(1, n // 8)
        t0 = time.perf_counter()
        ros2_c.ros2_steps(ys[:m1], ks[:m1], recs[:m1], 1e-10, net, nthreads=1)
        cpu_1 = m1 / (time.perf_counter() - t0)
        cpu = {"value": round(cpu_rate, 3), "unit": "column-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"{n} column-steps (one Ros2 step of {n} columns, nz=150, ni=93), "
                         f"LU-based C restatement, one column per OpenMP thread",
               "single_thread": round(cpu_1, 3), "host_cpus": ncpu}
        print(f"[bench] cpu: {cpu_rate:.2f} column-steps/s on {threads} threads, "
              f"{cpu_1:.2f} single-thread", file=sys.stderr, flush=True)

    metric = json.loads((ROOT / "BASELINE.json").read_text())["metric"]
    out = {
        "metric": metric,
        "value": round(value, 3),
        "unit": "column-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded: workload-1 distributions, one shared network)",
        "config": {"workload": "Ros2 step, nz=150 ni=93 S=2461 (1200 reversible + 60 photo rows), "
                               f"{B} independent columns per step, dt=1e-10",
                   "batch_columns": B, "nz": NZ, "ni": NI, "rate_rows": S},
        "factor_kernel": main_factor,
        "single_column_steps_per_s": round(single, 3),
        "single_column_cr_steps_per_s": round(single_cr, 3),
        "batched_over_single": round(value / single, 2),
        "batched_over_single_cr": round(value / single_cr, 2),
        "kernel_ms_per_launch": {"chem_lhs": round(kern_ms[0], 4), "bt_factor": round(kern_ms[1], 4),
                                 "bt_back1": round(kern_ms[2], 4), "rhs2": round(kern_ms[3], 4),
                                 "bt_solve2": round(kern_ms[4], 4),
                                 "step_prep": round(kern_ms[6], 4)},
        "step_tflops": round(B * step_flops(NZ, NI, S) / (elapsed / args.steps) / 1e12, 3),
        "finite": finite,
        "singular_columns": n_singular,
        "uneven_batch": uneven,
        "serving": serving,
        "config2_steady_state": config2,
        "config4_gradient": config4,
        "stopping_rule_cost": conv_cost,
        "solve_latency": solve_latency,
        "rate_build": rate_build,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if mlog is not None:
        for leg in ("uneven_batch", "serving", "config2_steady_state", "config4_gradient",
                    "rate_build", "roofline", "cpu_baseline"):
            if out.get(leg) is not None:
                mlog.write("bench_leg", leg=leg, **out[leg])
        mlog.write("bench_result", **{k: v for k, v in out.items()
                                      if not isinstance(v, dict) or k == "config"})
        mlog.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
