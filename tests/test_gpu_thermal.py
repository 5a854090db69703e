"""HIP conduction / transitions vs the oracle (bit-exact), interleaved with
the map rebuild the way the tick runs them."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu
FIELDS = ["part_ty", "part_life", "part_ctype", "part_temp", "part_equil", "part_tmp", "part_tmp2",
          "air_hv", "air_pv", "block_airh", "wall_emap", "pmap"]


def hot_scene(workload, scale):
    st = S.make_scene(workload, scale_particles=scale)
    rng = np.random.default_rng(11)
    live = np.flatnonzero(st.arrays["part_ty"][: st.slot_bound])
    # a spread of temperatures around the water/steam transitions
    st.arrays["part_temp"][live] = rng.uniform(250.0, 420.0, len(live)).astype(np.float32)
    st.arrays["air_hv"][:] = rng.uniform(280.0, 330.0, st.arrays["air_hv"].size).astype(np.float32)
    return st


@pytest.mark.parametrize("workload,scale,ticks", [("612x384", 1.0, 8), ("1600x900", 0.5, 4), ("2048x2048", 0.25, 3)])
def test_thermal_bit_exact(gpu, workload, scale, ticks):
    st = hot_scene(workload, scale)
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[thermal parity] {workload}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            seed = L.frame_seed(frame)
            ctx.recalc_free_particles(seed, False)
            o.call("oracle_recalc_free_particles", seed, 0)
            ctx.step_thermal(seed)
            o.call("oracle_step_thermal", seed, 0)
        ctx.fetch(st, FIELDS)
    for n in FIELDS:
        a, b = ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)
        assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} differ"


def test_legacy_mode_transitions_only(gpu):
    st = hot_scene("612x384", 0.5)
    ref = st.copy()
    o = O.OracleState(ref)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        ctx.recalc_free_particles(5, False)
        o.call("oracle_recalc_free_particles", 5, 0)
        ctx.step_thermal(5, legacy=True)
        o.call("oracle_step_thermal", 5, 1)
        ctx.fetch(st, FIELDS)
    for n in FIELDS:
        assert np.array_equal(ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)), n


@pytest.mark.parametrize("pixel_major", [True, False], ids=["pixel-major", "per-slot"])
def test_thermal_rare_branches(gpu, pixel_major):
    """Stacks, INVIS/FILT, photons, HSWC/PIPE/GEL/FILT neighbourhoods and
    off-pixel positions, through both conduction forms: right after
    RecalcFreeParticles the pixel-major form runs; an upload in between
    makes the library fall back to the per-slot form."""
    from edge_scenes import adversarial, off_pixel_positions, thermal_specials
    st = hot_scene("612x384", 0.5)
    rng = np.random.default_rng(23)
    adversarial(st, rng)
    thermal_specials(st, rng)
    off_pixel_positions(st, rng)
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[thermal rare branches] pixel_major={pixel_major}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(4):
            seed = L.frame_seed(frame)
            stack = frame % 2 == 0
            ctx.recalc_free_particles(seed, stack)
            o.call("oracle_recalc_free_particles", seed, 1 if stack else 0)
            if not pixel_major:
                ctx.upload("part_tmp2", ctx.download("part_tmp2", np.int32))
            ctx.step_thermal(seed)
            o.call("oracle_step_thermal", seed, 0)
        ctx.fetch(st, FIELDS)
    for n in FIELDS:
        a, b = ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)
        assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} differ"
