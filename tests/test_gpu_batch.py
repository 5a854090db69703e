"""Graph replay of the tick and the batched form, bit-exact against the
plain tpt_step loop (the frame counter, seed and CheckStacking flag come
from device memory in a replay)."""
import numpy as np
import pytest

from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu
FIELDS = [n for n in S.BUFFER_ORDER if n not in S.TABLES]


def same(a, b, label):
    for n in FIELDS:
        x, y = a.arrays[n].view(np.uint32), b.arrays[n].view(np.uint32)
        assert np.array_equal(x, y), f"{label}: {n} differs at {np.count_nonzero(x != y)} entries"


def test_run_replays_equal_the_step_loop(gpu):
    st = S.make_scene("612x384", scale_particles=0.5)
    a, b = st.copy(), st.copy()
    print("[graph] run vs step loop, 612x384", flush=True)
    with L.Context(st.cfg) as ca, L.Context(st.cfg) as cb:
        ca.load_state(a)
        cb.load_state(b)
        ca.run(0, 13)               # frames 0..12: frame 10 stacks from the device flag
        ca.run(13, 6)               # a second call reuses the graph
        for f in range(19):
            cb.step(f)
        ca.fetch(a)
        cb.fetch(b)
    same(a, b, "graph replay")


def test_batch_equals_single_canvases(gpu):
    shapes = [(612, 384, 0, 20000), (256, 128, 3, 3000), (1024, 512, 7, 60000)]
    scenes = [S.make_scene(xres=x, yres=y, seed=sd, particles=n) for x, y, sd, n in shapes]
    ref = [s.copy() for s in scenes]
    got = [s.copy() for s in scenes]
    print("[batch] 3 canvases vs single contexts", flush=True)
    with L.Batch([s.cfg for s in scenes]) as batch:
        for c, s in zip(batch.contexts, got):
            c.load_state(s)
        batch.step(0, 7)
        batch.step(7, 4)
        batch.synchronize()
        for c, s in zip(batch.contexts, got):
            c.fetch(s)
    for s in ref:
        with L.Context(s.cfg) as c:
            c.load_state(s)
            for f in range(11):
                c.step(f)
            c.fetch(s)
    for k, (r, g) in enumerate(zip(ref, got)):
        same(r, g, f"canvas {k}")


def test_batch_launches_equal_graph(gpu):
    shapes = [(400, 300, 1, 8000), (800, 200, 2, 12000)]
    scenes = [S.make_scene(xres=x, yres=y, seed=sd, particles=n) for x, y, sd, n in shapes]
    a = [s.copy() for s in scenes]
    b = [s.copy() for s in scenes]
    print("[batch] launches vs graph", flush=True)
    with L.Batch([s.cfg for s in scenes]) as ba, L.Batch([s.cfg for s in scenes]) as bb:
        for c, s in zip(ba.contexts, a):
            c.load_state(s)
        for c, s in zip(bb.contexts, b):
            c.load_state(s)
        ba.step(0, 9)
        bb.step_launches(0, 9)
        ba.synchronize()
        bb.synchronize()
        for c, s in zip(ba.contexts, a):
            c.fetch(s)
        for c, s in zip(bb.contexts, b):
            c.fetch(s)
    for k in range(len(scenes)):
        same(a[k], b[k], f"canvas {k}")


def run_batch(scenes, drive):
    """Load `scenes` into a fresh batch, apply drive(batch), return the states."""
    out = [s.copy() for s in scenes]
    with L.Batch([s.cfg for s in scenes]) as b:
        for c, s in zip(b.contexts, out):
            c.load_state(s)
        drive(b)
        b.synchronize()
        for c, s in zip(b.contexts, out):
            c.fetch(s)
    return out


def scenes_for_lanes():
    shapes = [(400, 300, 1, 8000), (800, 200, 2, 12000), (256, 256, 5, 5000), (128, 512, 6, 4000), (640, 480, 8, 30000)]
    return [S.make_scene(xres=x, yres=y, seed=sd, particles=n) for x, y, sd, n in shapes]


def test_serial_and_lanes_equal_launches(gpu):
    """One canvas at a time and the whole-batch graph split into 2, 3 and one
    lane per canvas give the
    per-canvas launches' results."""
    scenes = scenes_for_lanes()
    print("[batch] serial / lanes vs launches", flush=True)
    ref = run_batch(scenes, lambda b: b.step_launches(0, 12))
    got = run_batch(scenes, lambda b: b.step_serial(0, 12))
    for k in range(len(scenes)):
        same(ref[k], got[k], f"serial canvas {k}")
    for lanes in (2, 3, len(scenes)):
        def drive(b, lanes=lanes):
            b.set_lanes(lanes)
            b.step(0, 12)
        got = run_batch(scenes, drive)
        for k in range(len(scenes)):
            same(ref[k], got[k], f"{lanes} lanes canvas {k}")


@pytest.mark.skip(reason="tpt_batch_autotune is switched off: one run aborted inside the HIP runtime "
                         "(no GPU fault reported) with no cause found")
def test_autotune_equals_launches(gpu):
    """The autotuned graph gives the per-canvas launches' results."""
    scenes = scenes_for_lanes()
    ref = run_batch(scenes, lambda b: b.step_launches(0, 12))
    seen = {}

    def tuned(b):
        best, table = b.autotune(0, 1)            # 1 eager tick + 2 replays per candidate
        seen["best"], seen["table"] = best, table
        done = 1 + 2 * len(table)
        b.step(done, 12 - done)
    got = run_batch(scenes, tuned)
    assert [t[0] for t in seen["table"]] == [1, 2, 4, 0] and seen["best"] in (1, 2, 4, 0)
    assert all(ms > 0 for _, ms in seen["table"])
    for k in range(len(scenes)):
        same(ref[k], got[k], f"autotuned canvas {k}")


def test_autotune_is_switched_off():
    """The switched-off autotuner refuses with TPT_ESTATE (no GPU work)."""
    from tpt_hip import scene as S2
    with L.Batch([S2.default_config(64, 64)]) as b:
        with pytest.raises(L.TptError):
            b.autotune(0, 1)
