"""Whole ticks under each simulation mode the config carries, HIP vs the
oracle bit for bit on the 612x384 workload (half its particles): the edge
modes without the wall ring (so that particles reach the border), the
gravity modes including radial and a custom vector, the convection modes,
the legacy heat path, non-zero edge pressure and velocity, no vorticity,
and the air modes inside whole ticks."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import lib as L
from tpt_hip import scene as S
from test_gpu_move import FLOAT_FIELDS, INT_FIELDS, compare

pytestmark = pytest.mark.gpu
FIELDS = INT_FIELDS + FLOAT_FIELDS + ["pmap_count", "block_air", "block_airh", "wall_emap"]

VARIANTS = {
    "edge_solid": {"edge_mode": 1},
    "edge_loop": {"edge_mode": 2},
    "edge_void": {"edge_mode": 0},
    "gravity_off": {"gravity_mode": 1},
    "gravity_radial": {"gravity_mode": 2},
    "gravity_custom": {"gravity_mode": 3, "custom_gravity_x": 0.3, "custom_gravity_y": -0.5},
    "convection_off": {"convection_mode": 0},
    "convection_legacy": {"convection_mode": 1},
    "heat_legacy": {"heat_enable": 0},
    "edge_air": {"edge_pressure": 5.0, "edge_velocity_x": 1.0, "edge_velocity_y": -0.5, "vorticity_coeff": 0.0},
    "air_pressure_off": {"air_mode": 1},
    "air_velocity_off": {"air_mode": 2},
    "air_off": {"air_mode": 3},
    "air_no_update": {"air_mode": 4},
}
OPEN_EDGES = {"edge_solid", "edge_loop", "edge_void", "gravity_radial", "gravity_custom"}


def scene(name):
    st = S.make_scene("612x384", scale_particles=0.5, newton_grav=False)
    if name in OPEN_EDGES:
        st.arrays["wall_bmap"][:] = 0
        S.derive_block_masks(st)
    for k, v in VARIANTS[name].items():
        st.cfg[k] = v
    return st


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_mode_whole_ticks_bit_exact(gpu, name):
    st = scene(name)
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[mode parity] {name}, 10 ticks", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(10):
            ctx.step(frame)
            o.call("oracle_step", frame)
        ctx.fetch(st)
    compare(ref, st, FIELDS, name)
    assert st.slot_bound == ref.slot_bound
    if name in ("edge_void", "edge_loop", "edge_solid"):
        # the border was reached: void kills, loop wraps, solid keeps
        assert np.count_nonzero(ref.arrays["part_ty"]) > 0
