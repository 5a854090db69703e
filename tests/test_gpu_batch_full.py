"""The batched form at its full shape: 256 canvases with sides drawn from
128..2048 in steps of 4 (seed 2), fills from 10-50%, the 612x384 workload's
element mix.  Five ticks of tpt_batch_step (one eager tick, then replays of
the graph holding every canvas) must equal five per-canvas tpt_step ticks of
the same canvases, bit for bit (floats by value)."""
import numpy as np
import pytest

from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu
TICKS = 5
SLOT = [("part_ty", np.uint32), ("part_x", np.float32), ("part_y", np.float32), ("part_vx", np.float32),
        ("part_vy", np.float32), ("part_temp", np.float32), ("part_life", np.int32), ("part_ctype", np.uint32),
        ("part_tmp", np.int32), ("part_flags", np.uint32)]
MAPS = [("pmap", np.uint32), ("photons", np.uint32), ("air_pv", np.float32), ("air_vx", np.float32),
        ("air_vy", np.float32), ("air_hv", np.float32)]


def config3_canvases(n=256):
    """(width, height, seed, particles) of the batched workload (bench.py --batched)."""
    rng = np.random.default_rng(2)
    shapes = [(4 * int(rng.integers(32, 513)), 4 * int(rng.integers(32, 513))) for _ in range(n)]
    fills = rng.uniform(0.10, 0.50, n)
    return [(w, h, 1000 + k, int(fills[k] * w * h)) for k, (w, h) in enumerate(shapes)]


def snapshot(ctx):
    bound = ctx.slot_bound()
    out = {n: ctx.download(n, dt, bound) for n, dt in SLOT}
    out.update({n: ctx.download(n, dt) for n, dt in MAPS})
    out["slot_bound"] = bound
    return out


def test_256_canvas_batch_equals_per_canvas_ticks(gpu):
    canvases = config3_canvases()
    print(f"[batch full] generating {len(canvases)} canvases", flush=True)
    states = [S.make_scene(xres=w, yres=h, seed=sd, particles=n) for (w, h, sd, n) in canvases]
    px = sum(w * h for (w, h, _, _) in canvases)
    print(f"[batch full] {px} pixels, {sum(st.slot_bound for st in states)} slots", flush=True)
    with L.Batch([st.cfg for st in states]) as b:
        for ctx, st in zip(b.contexts, states):
            ctx.load_state(st)
        b.step_launches(0, TICKS)
        b.synchronize()
        ref = [snapshot(ctx) for ctx in b.contexts]
        print("[batch full] per-canvas launches done", flush=True)
        for ctx, st in zip(b.contexts, states):
            ctx.load_state(st)
        b.step(0, TICKS)
        b.synchronize()
        print("[batch full] batched ticks done", flush=True)
        for k, ctx in enumerate(b.contexts):
            got = snapshot(ctx)
            assert got["slot_bound"] == ref[k]["slot_bound"], k
            for name, _ in SLOT + MAPS:
                assert np.array_equal(got[name], ref[k][name]), (k, canvases[k], name)
    moved = sum(int(np.count_nonzero(r["part_y"] != st.arrays["part_y"][: r["slot_bound"]])) for r, st in zip(ref, states))
    assert moved > 0
