"""The three synthetic workloads at their full sizes, whole ticks, HIP vs
the oracle bit for bit:

* 612x384, 20,000 particles, 1000 ticks through ``tpt_run`` (graph
  replays), compared at ticks 250, 500 and 1000 — the run length the
  612x384 workload is specified with;
* 1600x900, ~403,000 particles with the FIRE and OIL patches, 40 ticks;
* 2048x2048, 2,097,152 particles and 64 wall rectangles, 6 ticks.

The element mixes carry no Newtonian mass under the reference's mass
table, so the force grids are exact zeros and every field is compared
exactly."""
import pytest

from oracle import oracle as O
from tpt_hip import lib as L
from tpt_hip import scene as S
from test_gpu_move import FLOAT_FIELDS, INT_FIELDS, compare

pytestmark = pytest.mark.gpu
FIELDS = INT_FIELDS + FLOAT_FIELDS + ["pmap_count", "block_air", "block_airh", "wall_emap", "grav_mass",
                                      "grav_mass_prev", "grav_fx", "grav_fy"]


def test_612x384_1000_ticks_graph_replay(gpu):
    st = S.make_scene("612x384")
    ref = st.copy()
    o = O.OracleState(ref)
    frame = 0
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for stop in (250, 500, 1000):
            n = stop - frame
            ctx.run(frame, n)
            for f in range(frame, stop):
                o.call("oracle_step", f)
            frame = stop
            ctx.fetch(st)
            print(f"[config parity] 612x384 tick {stop}", flush=True)
            compare(ref, st, FIELDS, f"612x384 tick {stop}")
            assert st.slot_bound == ref.slot_bound


@pytest.mark.parametrize("workload,ticks", [("1600x900", 40), ("2048x2048", 6)])
def test_full_size_whole_ticks(gpu, workload, ticks):
    st = S.make_scene(workload)
    ref = st.copy()
    o = O.OracleState(ref)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            ctx.step(frame)
            o.call("oracle_step", frame)
            if frame % 10 == 9:
                print(f"[config parity] {workload} tick {frame + 1}", flush=True)
        ctx.fetch(st)
    compare(ref, st, FIELDS, f"{workload} tick {ticks}")
    assert st.slot_bound == ref.slot_bound
