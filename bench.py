"""Benchmark: plain HIP tick throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 2048x2048]
                    [--water-equal] [--scene SAVE.tptrs] [--batched N]

Prints ONE JSON line.  A step is one whole tick over the synthetic
2048x2048 workload (512x512 air cells, ~2.1M particles with FIRE/OIL
patches, 64 wall rectangles), with every buffer resident in HBM before
timing starts.  `value` times the tick through tpt_run (captured once as a
hipGraph, then replayed); `step_loop` times the same frames through one
tpt_step call per tick.  `roofline` reports the dominant kernel's achieved
bandwidth from HIP events recorded on the library's own stream over the
same frames; `cpu_baseline` times the CPU oracle
on a bounded number of ticks of the same workload (all threads, and per
workload on all threads and one).  `batched` is the 256-canvas form (graph
replay (one chain per canvas) against one canvas at a time and against
per-canvas stream launches); `surrogate` is the learned air step against 8
plain air ticks on the 512x512 air grid.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "the-powder-toy-rust-wgpu_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0

# The whole tick: tpt_step (SimulationEngine::step order, lib.rs:334-409).
TICK_PASSES = ["tick"]


def kernel_bytes(name, C, P, S, N, W, yc):
    """Algorithmic bytes of one launch of each timed pass: every array the pass
    must stream, once (stencil re-reads, LDS reuse and cache hits are free).
    C = air cells, P = pixels, S = scanned slots, N = live particles,
    W = padded FFT row length, yc = air-cell rows."""
    table = {
        "air_pressure_velocity": 28 * C,      # pv,vx,vy,block_air in; opv,ovx,ovy out
        "air_smooth_advect": 40 * C,          # ovx,ovy,opv,block_air,wall,fan x2 in; vx,vy,pv out
        "air_heat_convect": 28 * C,           # hv,vx,vy,block_airh in; ohv,vx',vy' out
        "air_settle": 124 * C,                # 9 grids + 16 pmap words in; 7 grids out
        "grav_fft_rows": 8 * C + 8 * yc * W,  # mass,mask in; ycells padded rows out
        "grav_fft_cols": None,                # filled in below (needs H)
        "grav_fft_inverse_extract": 16 * yc * W + 12 * C,
        "grav_apply": 4 * S + 36 * N,         # ty scan; x,y,vx,vy,flags,force gather in; vx,vy out; mass deposit
        "grav_build_mass": 8 * C,             # mass -> snapshot copy (element-weight mass mode only)
        "recalc_clear": 8 * P,                # pmap_count from the movement rebuild's counts
        "recalc_scatter": 4 * S + 12 * P,     # ty scan (life decay); records in, pmap out
        "check_stacking": 4 * P + 4 * S + 12 * N,
        "thermal_conduction": 16 * P + 4 * N + 16 * C,  # records in, per-pixel results out; equil out
        "thermal_conduction_cells": 16 * C,
        "thermal_apply": 16 * P + 8 * N,      # records + results in; temp out, record temp out
        "phase_transition": 12 * S + 72 * N,  # convection (4 words + partner record) + transitions
        "phase_transition_cells": 16 * C,
        "refresh_pmap_types": 4 * S + 16 * N,
        "air_deposit": 8 * P + 16 * C,        # records in; per-cell winner gathers; air_vx,vy RW
        "move_prepare": 9 * P + 32 * C,       # cell_winner clear; pmap in, type bytes out; cell records
        "move_velocity_intent": 4 * S + 52 * N,  # ty scan; x,y,vx,vy,flags,tmp,air in; vx,vy,intent out; claim
        "move_saw_pair": 0,
        "move_commit": 4 * S + 32 * N,
        "move_rebuild_maps": 36 * P + 4 * S + 20 * N,  # records/photons/counts clear; scatter; pmap + counts
    }
    return table.get(name)


# Timed passes that are one kernel, for the roofline's HBM traffic (the
# rocprofv3 summary is per kernel).
PASS_KERNEL = {"move_velocity_intent": "mv_intent", "move_commit": "mv_commit", "grav_apply": "grav_apply_mass",
               "grav_fft_cols": "grav_cols", "thermal_apply": "th_apply_px"}


def reference_tick_bytes(C, P, N, G):
    """Algorithmic bytes of one tick of the reference's passes, every logical
    array read once and written once per pass (C air cells, P pixels, N live
    particles, G cells of the padded FFT grid): air 140C; gravity FFT and
    mass 96G + 32C + 36N; conduction and apply 32N + 4P + 16C; movement
    68N + 20P + 24C; RecalcFreeParticles 12P + 32N; phase transitions and
    the pmap refresh 68N + 28C."""
    return (140 * C + (96 * G + 32 * C + 36 * N) + (32 * N + 4 * P + 16 * C) + (68 * N + 20 * P + 24 * C)
            + (12 * P + 32 * N) + (68 * N + 28 * C))


def traffic_table(workload):
    """Calibrated HBM bytes per dispatch by kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, if any."""
    path = os.path.join(ROOT, "profiles", f"tick_{workload}_traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        return json.load(fh)["kernels"], os.path.relpath(path, ROOT)


def run_pass(ctx, name, frame):
    from tpt_hip import lib as L
    seed = L.frame_seed(frame)
    if name == "air":
        ctx.step_air()
    elif name == "gravity":
        ctx.step_gravity(seed)
    elif name == "recalc":
        ctx.recalc_free_particles(seed, frame % 10 == 0)
    elif name == "thermal":
        ctx.step_thermal(seed)
    elif name == "movement":
        ctx.step_movement(seed)
    elif name == "tick":
        ctx.step(frame)


def oracle_pass(o, name, frame):
    if name == "air":
        o.call("oracle_step_air")
    elif name == "tick":
        o.call("oracle_step", frame)


def batched_leg(n, steps, warmup, log=None):
    """The batched form: n independent canvases with sides drawn from
    128..2048 in steps of 4 (seed 2), each with its own seed, a fill drawn
    from 10-50%, the 612x384 workload's element mix and init rules.  Times
    ticks four ways: one canvas at a time (the reference steps one engine
    after another), per-canvas tpt_step launches on their own streams,
    tpt_batch_step (each canvas's captured tick replayed on its own stream;
    the `value`), and the whole-batch graph with a chain per canvas.  The
    autotuner that picks a work split of the whole-batch graph is switched
    off."""
    import numpy as np
    from tpt_hip import lib as L
    from tpt_hip import scene as S

    rng = np.random.default_rng(2)
    shapes = [(4 * int(rng.integers(32, 513)), 4 * int(rng.integers(32, 513))) for _ in range(n)]
    fills = rng.uniform(0.10, 0.50, n)
    t0 = time.perf_counter()
    batch = L.Batch([S.default_config(w, h) for w, h in shapes])
    particles = 0
    for k, (w, h) in enumerate(shapes):
        st = S.make_scene(xres=w, yres=h, seed=1000 + k, particles=int(fills[k] * w * h))
        batch.contexts[k].load_state(st)
        particles += int((st.arrays["part_ty"] != 0).sum())
        del st
        if log and k % 64 == 63:
            log(f"[batched] {k + 1}/{n} canvases loaded ({time.perf_counter() - t0:.0f} s)")
    frame = 0

    def timed(fn, ticks):
        nonlocal frame
        batch.synchronize()
        t = time.perf_counter()
        fn(frame, ticks)
        batch.synchronize()
        frame += ticks
        return time.perf_counter() - t

    timed(batch.step_launches, warmup)
    t_serial = timed(batch.step_serial, steps)
    t_launch = timed(batch.step_launches, steps)
    timed(batch.step, 2)  # eager tick + one replay: the graphs are captured
    t_graph = timed(batch.step, steps + 1) * steps / (steps + 1)  # its eager first tick included
    batch.set_lanes(n)    # one graph for the whole batch, a chain per canvas
    timed(batch.step, 2)
    t_whole = timed(batch.step, steps + 1) * steps / (steps + 1)
    batch.close()
    px = sum(w * h for w, h in shapes)
    rate = lambda t: round(n * steps / t, 2)  # noqa: E731
    return {
        "value": rate(t_graph), "unit": "canvas-ticks/s", "ms_per_batch_tick": round(t_graph / steps * 1e3, 3),
        "workload": f"{n} canvases, sides 128..2048 step 4 (seed 2), fill 10-50%",
        "pixels": px, "particles": particles,
        "serial_one_canvas_at_a_time": {"value": rate(t_serial), "ms_per_batch_tick": round(t_serial / steps * 1e3, 3)},
        "per_canvas_stream_launches": {"value": rate(t_launch), "ms_per_batch_tick": round(t_launch / steps * 1e3, 3)},
        "whole_batch_graph": {"value": rate(t_whole), "ms_per_batch_tick": round(t_whole / steps * 1e3, 3),
                              "note": "one hipGraph holding every canvas's tick as a parallel chain"},
        "speedup_vs_serial": round(t_serial / t_graph, 3),
        "speedup_vs_stream_launches": round(t_launch / t_graph, 3),
        "graphs": "each canvas's tick captured as its own hipGraph, replayed on the canvas's stream",
    }


def batched(args):
    out = batched_leg(args.batched, args.steps, args.warmup,
                      log=lambda m: print(m, file=sys.stderr, flush=True))
    out.update({"metric": "aggregate canvas-ticks/s, batched form", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic"})
    print(json.dumps(out))


# Algorithmic FLOPs of one surrogate call: 2 * pixels * 9 * cin * cout per
# 3x3 convolution of the U-Net (real channels, 512x512 grid).
SURROGATE_LAYERS = [(0, 6, 32), (1, 32, 64), (1, 64, 64), (2, 64, 128), (2, 128, 128), (2, 128, 64),
                    (1, 64, 64), (1, 64, 32), (0, 32, 32), (0, 32, 4)]
MFMA_BF16_PEAK_TFLOPS = 2500.0


def surrogate_leg(seeds=16):
    """The learned air step on the 512x512 air grid: call time against 8
    plain air ticks, relative L2 against the plain solver after 64 ticks on
    held-out scenes, fallback rate, MFMA utilisation."""
    from tpt_hip import lib as L
    from tpt_hip import scene as S
    from tpt_hip import surrogate as SU

    if not os.path.exists(SU.DEFAULT_WEIGHTS):
        return None
    xc = 512
    cfg = SU.air_config(xc, xc)
    ctx = L.Context(cfg)
    ctx.load_state(S.SimState(cfg))
    sur = SU.Surrogate(xc, xc)
    meta = sur.load(SU.DEFAULT_WEIGHTS)
    t = SU.timing(sur, ctx, reps=50)
    ev = SU.evaluate(sur, ctx, range(10 ** 7, 10 ** 7 + seeds))
    ctx.close()
    sur.close()
    flops = sum(2 * (xc >> lv) ** 2 * 9 * ci * co for lv, ci, co in SURROGATE_LAYERS)
    tf = flops / (t["surrogate_call"] * 1e-6) / 1e12
    return {
        "workload": "512x512 air grid, 8 ticks per call, held-out scenes (4-16 blobs, 0-32 walls, 0-8 heat sources)",
        "call_us": round(t["surrogate_call"], 1), "call_checked_us": round(t["surrogate_call_checked"], 1),
        "plain_8_ticks_us": round(t["plain_8_ticks"], 1),
        "speedup_vs_plain": round(t["plain_8_ticks"] / t["surrogate_call"], 3),
        "speedup_vs_plain_checked": round(t["plain_8_ticks"] / t["surrogate_call_checked"], 3),
        "rel_l2_64_ticks": ev["rel_l2"], "rel_l2_8_ticks": ev["rel_l2_8_ticks"], "eval_seeds": seeds,
        "fallback_rate": ev["fallback_rate"],
        "roofline": {"bound": "mfma", "achieved": round(tf, 1), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tf / MFMA_BF16_PEAK_TFLOPS, 4), "flops_per_call": flops},
        "trained_epochs": meta.get("epochs"), "train_trajectories": meta.get("traj"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=int(os.environ.get("WORLD_SIZE", "1")))
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="2048x2048")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--batched", type=int, default=0, metavar="N",
                    help="the batched form alone: N mixed-size canvases")
    ap.add_argument("--batch-canvases", type=int, default=256,
                    help="canvases of the batched leg of the default run (0 skips it)")
    ap.add_argument("--no-surrogate", action="store_true", help="skip the surrogate leg")
    ap.add_argument("--water-equal", action="store_true",
                    help="run the workload with water_equal_test set (the optional pass after movement)")
    ap.add_argument("--scene", default=None, metavar="PATH",
                    help="tick a .tptrs save instead of the synthetic workload")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.empty(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(10):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": a.numel() * 4, "launches": 10}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.batched:
        return batched(args)

    from tpt_hip import lib as L
    from tpt_hip import scene as S

    def scene():
        if args.scene:
            from tpt_hip import save as V
            return V.restore(V.load(args.scene))
        return S.make_scene(args.workload)

    st = scene()
    if args.scene:
        args.workload = os.path.basename(args.scene)
    if args.water_equal:
        st.cfg["water_equal_test"] = 1
    C, P = st.xcells * st.ycells, st.cfg["xres"] * st.cfg["yres"]
    N = int((st.arrays["part_ty"] != 0).sum())
    S_ = st.slot_bound
    W = 1
    while W < 2 * st.xcells:
        W *= 2
    H = 1
    while H < 2 * st.ycells:
        H *= 2
    yc = st.ycells
    passes = TICK_PASSES

    def eager(ctx, first, n):
        for f in range(first, first + n):
            for p in passes:
                run_pass(ctx, p, f)

    # Every leg ticks frames 0 .. warmup-1 untimed, then times frames
    # warmup .. warmup+steps-1 of the same initial state, so that the three
    # legs do identical work.
    # (1) tpt_run: the whole tick captured once as a hipGraph and replayed
    #     (the `value`).
    ctx = L.Context(st.cfg)
    ctx.load_state(st)
    ctx.run(0, args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.run(args.warmup, args.steps)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    ticks_per_s = args.steps / dt
    # (2) the tpt_step loop: one host call per tick, every kernel launched
    #     from the host.
    ctx.load_state(st)
    eager(ctx, 0, args.warmup)
    ctx.synchronize()
    t0 = time.perf_counter()
    eager(ctx, args.warmup, args.steps)
    ctx.synchronize()
    dt_eager = time.perf_counter() - t0
    # (3) per-pass durations from HIP events on the library's stream, over
    #     the same frames ticked once more through tpt_step.
    ctx.load_state(st)
    eager(ctx, 0, args.warmup)
    ctx.set_profiling(True)
    ctx.reset_pass_times()
    eager(ctx, args.warmup, args.steps)
    times = ctx.pass_times()
    ctx.set_profiling(False)
    per_kernel = {}
    for name, (ms, n) in times.items():
        if n:
            b = kernel_bytes(name, C, P, S_, N, W, yc)
            if name == "grav_fft_cols":
                b = 8 * yc * W * 3 + 16 * W * H   # work in, 2 spectra out, 2 kernel spectra in
            avg_us = ms / n * 1e3
            per_kernel[name] = {"avg_us": round(avg_us, 3), "launches": n,
                                "alg_bytes": b,
                                "gbs": round(b / (avg_us * 1e-6) / 1e9, 1) if b else None}
    # The roofline line is for one kernel: the single-kernel pass with the
    # largest share of the tick.
    dom = max((k for k in per_kernel if k in PASS_KERNEL),
              key=lambda k: per_kernel[k]["avg_us"] * per_kernel[k]["launches"])
    d = per_kernel[dom]
    achieved = d["gbs"]
    traffic, traffic_src = traffic_table(args.workload)
    kern = PASS_KERNEL.get(dom)
    measured = traffic.get(kern, {}).get("hbm_bytes") if (traffic and kern) else None
    # The whole tick against the same roofline: the algorithmic bytes of
    # every timed pass with a byte model, per tick, over the timed tick.
    design_bytes = sum(v["alg_bytes"] * v["launches"] for v in per_kernel.values() if v["alg_bytes"]) / args.steps
    tick_bytes = reference_tick_bytes(C, P, N, W * H)
    tick_gbs = tick_bytes / (dt / args.steps) / 1e9
    tick_roofline = {"alg_bytes_per_tick": round(tick_bytes), "achieved": round(tick_gbs, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(tick_gbs / HBM_PEAK_GBS, 4),
                     "byte_model": "reference passes, each array streamed once: air 140C, gravity 96G+32C+36N, "
                                   "conduction 32N+4P+16C, movement 68N+20P+24C, recalc 12P+32N, "
                                   "transitions 68N+28C",
                     "design_bytes_per_tick": round(design_bytes),
                     "design_frac": round(design_bytes / (dt / args.steps) / 1e9 / HBM_PEAK_GBS, 4)}
    roofline = {"bound": "hbm", "kernel": kern or dom, "pass": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                "traffic": measured, "traffic_source": traffic_src if measured else None,
                "alg_bytes_per_launch": d["alg_bytes"], "avg_us": d["avg_us"]}
    ctx.close()

    cpu, cpu_configs = None, None
    if not args.no_cpu:
        from oracle import oracle as O
        olib = O.lib()
        all_threads = olib.oracle_max_threads()

        def oracle_rate(state, threads, seconds, max_ticks):
            """Ticks/s of the C oracle over a bounded sample of whole ticks."""
            olib.oracle_set_threads(threads)
            o = O.OracleState(state)
            n_cpu, t0 = 0, time.perf_counter()
            while True:
                for p in passes:
                    oracle_pass(o, p, n_cpu)
                n_cpu += 1
                el = time.perf_counter() - t0
                if el >= seconds or n_cpu >= max_ticks:
                    return n_cpu, n_cpu / el

        ref = scene()
        ref.cfg["water_equal_test"] = st.cfg["water_equal_test"]
        n_cpu, rate = oracle_rate(ref, all_threads, args.cpu_seconds, args.steps)
        cpu = {"value": round(rate, 3), "unit": "ticks/s", "cores": all_threads, "kind": "port",
               "sample": f"{n_cpu} ticks of the same {args.workload} workload and passes, C oracle, "
                         f"OpenMP over {all_threads} threads (air stencils, FFT rows, gravity and "
                         f"thermal apply, conduction; the passes with ordered rules serial)"}
        # Configs 0-2 on all threads and on one, a few seconds each.
        cpu_configs = {}
        if not args.scene:
            for wl in ("612x384", "1600x900", "2048x2048"):
                row = {}
                for threads in (all_threads, 1):
                    state = S.make_scene(wl)
                    n_cpu, rate = oracle_rate(state, threads, args.cpu_seconds / 3.0, 1000)
                    row[f"threads_{threads}"] = {"ticks_per_s": round(rate, 3), "ticks": n_cpu}
                cpu_configs[wl] = row
        olib.oracle_set_threads(all_threads)

    batched_out = None
    if args.batch_canvases and not args.scene:
        batched_out = batched_leg(args.batch_canvases, min(args.steps, 10), 2,
                                  log=lambda m: print(m, file=sys.stderr, flush=True))
    surrogate_out = None if args.no_surrogate else surrogate_leg()

    out = {
        "metric": "ticks/s @2048x2048 plain vs batched vs air-surrogate (relL2<=1e-2); HBM GB/s",
        "value": round(ticks_per_s, 2),
        "unit": "ticks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "saved scene" if args.scene else "synthetic",
        "config": {"workload": f"{args.workload} plain tick{' with water equalisation' if args.water_equal else ''}, "
                               f"passes: {'+'.join(passes)}, frames {args.warmup}..{args.warmup + args.steps - 1} "
                               f"through tpt_run (hipGraph replay)",
                   "air_cells": C, "pixels": P, "particles": N, "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_configs": cpu_configs,
        "step_loop": {"value": round(args.steps / dt_eager, 2), "ms_per_step": round(dt_eager / args.steps * 1e3, 4),
                      "note": "the same frames through one tpt_step call per tick (kernels launched from the host)"},
        "tick_roofline": tick_roofline,
        "per_kernel": per_kernel,
        "batched": batched_out,
        "surrogate": surrogate_out,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
