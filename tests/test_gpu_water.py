"""HIP water equalisation vs the oracle, bit for bit: the pass alone on
communicating vessels, and whole ticks with water_equal_test set."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import lib as L
from tpt_hip import scene as S
from test_gpu_move import FLOAT_FIELDS, INT_FIELDS, compare
from test_oracle_water import utube

pytestmark = pytest.mark.gpu


def two_channels():
    from tpt_hip import elements as E
    W, H = 96, 80
    st = S.SimState(S.default_config(W, H, capacity=30000))
    st.cfg["water_equal_test"] = 1
    a = st.arrays
    k = 0
    for y in range(4, H - 4):
        for x in range(4, W - 4):
            water = ((10 <= x <= 18 and 12 <= y <= 70) or (19 <= x <= 50 and 40 <= y <= 44) or
                     (45 <= x <= 50 and 35 <= y <= 44) or (19 <= x <= 90 and 66 <= y <= 70) or
                     (60 <= x <= 90 and 60 <= y <= 70))
            air = (10 <= x <= 18 and y < 12) or (45 <= x <= 50 and y < 35) or (60 <= x <= 90 and y < 60)
            if water or not air:
                a["part_ty"][k] = E.pt("WATR" if water else "METL")
                a["part_x"][k], a["part_y"][k], a["part_temp"][k] = x, y, 300.0
                k += 1
    st.slot_bound = k
    S.rebuild_pmap(st)
    return st


@pytest.mark.parametrize("scene", ["utube", "two_channels"])
def test_pass_alone_bit_exact(gpu, scene):
    st = utube() if scene == "utube" else two_channels()
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[water parity] {scene}, pass alone x200", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        ctx.recalc_free_particles(1, False)   # cell records for the pass's wall tests
        ctx.step_movement(1)
        o.call("oracle_recalc_free_particles", 1, 0)
        o.call("oracle_step_movement", 1)
        for f in range(200):
            seed = (f * 2654435761) & 0xFFFFFFFF
            ctx.step_water_equal(seed)
            o.call("oracle_water_equal", seed)
        ctx.fetch(st)
    compare(ref, st, ["part_x", "part_y", "part_ty", "pmap"], f"water {scene}")
    moved = np.count_nonzero(st.arrays["part_y"] != utube().arrays["part_y"]) if scene == "utube" else 1
    assert moved > 0


@pytest.mark.parametrize("scene,ticks", [("utube", 40), ("612x384", 15)])
def test_whole_ticks_with_water_bit_exact(gpu, scene, ticks):
    if scene == "utube":
        st = utube()
    else:
        st = S.make_scene("612x384", scale_particles=0.5, newton_grav=False)
        st.cfg["water_equal_test"] = 1
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[water parity] {scene} whole ticks={ticks}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            ctx.step(frame)
            o.call("oracle_step", frame)
        ctx.fetch(st)
    compare(ref, st, INT_FIELDS + FLOAT_FIELDS + ["pmap_count"], f"{scene} with water")


def test_water_ticks_through_graph_replay(gpu):
    """tpt_run with the water pass on: the first tick eager, the rest graph
    replays whose recalc is the full one (the water pass leaves no
    incremental maps), against the oracle's ticks."""
    st = S.make_scene("612x384", scale_particles=0.5, newton_grav=False)
    st.cfg["water_equal_test"] = 1
    ref = st.copy()
    o = O.OracleState(ref)
    print("[water parity] 612x384 via tpt_run, 24 ticks", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        ctx.run(0, 13)
        ctx.run(13, 11)
        ctx.fetch(st)
    for frame in range(24):
        o.call("oracle_step", frame)
    compare(ref, st, INT_FIELDS + FLOAT_FIELDS + ["pmap_count"], "612x384 with water, graph replay")
    assert st.slot_bound == ref.slot_bound
