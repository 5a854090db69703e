"""Whole ticks at the canvas extremes, HIP vs the oracle bit for bit: the
largest canvas the layout allows (2896x2896, capacity 0x7FFFFF), tiny and
empty canvases, a canvas whose cell grid is neither square nor a power of
two, and a slot range full of holes.  Gravity runs with the reference's
mass table (no mass in these mixes), so the force grids are exact zeros;
the FFT at these sizes is held to its tolerance in test_gpu_gravity."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import lib as L
from tpt_hip import scene as S
from test_gpu_move import FLOAT_FIELDS, INT_FIELDS, compare

pytestmark = pytest.mark.gpu
FIELDS = INT_FIELDS + FLOAT_FIELDS + ["pmap_count", "block_air", "block_airh", "wall_emap", "grav_mass", "grav_fx",
                                      "grav_fy"]


def run_ticks(st, ticks, label):
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[edge parity] {label} ticks={ticks}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            ctx.step(frame)
            o.call("oracle_step", frame)
        ctx.fetch(st)
    compare(ref, st, FIELDS, label)
    assert st.slot_bound == ref.slot_bound
    return ref


def test_largest_canvas(gpu):
    st = S.make_scene("custom", xres=2896, yres=2896, seed=5, particles=80000)
    assert st.cfg["capacity"] == 0x7FFFFF
    run_ticks(st, 2, "2896x2896")


@pytest.mark.parametrize("xres,yres,particles", [(16, 16, 20), (8, 8, 0), (4, 4, 0), (100, 36, 150)])
def test_small_and_empty_canvases(gpu, xres, yres, particles):
    st = S.make_scene("custom", xres=xres, yres=yres, seed=9, particles=particles)
    run_ticks(st, 12, f"{xres}x{yres} n={particles}")


def test_slot_range_with_holes(gpu):
    """Every third slot emptied after placement: the passes scan a slot
    range that is two-thirds live, with the window and the bound unchanged."""
    st = S.make_scene("612x384", scale_particles=0.5)
    ty = st.arrays["part_ty"]
    holes = np.arange(0, st.slot_bound, 3)
    ty[holes] = 0
    S.rebuild_pmap(st)
    run_ticks(st, 6, "612x384 with holes")
