"""Scenes that come in through the .tptrs loader, ticked by the HIP library
and by the oracle, bit for bit.  A loaded scene differs from the synthetic
ones: its particles fill consecutive slots from 0 (through the device spawn
window), positions are fractional and rounded by the map rebuild, pixels can
hold stacks and INVIS/FILT overlaps, energy particles share pixels with
matter, and the config is the loader's (Newtonian gravity off)."""
import numpy as np
import pytest

from tpt_hip import elements as E
from tpt_hip import save as V
from tpt_hip import scene as S
from test_gpu_edges import run_ticks

pytestmark = pytest.mark.gpu


def loaded(st, include_pressure=True, mutate=None):
    d = V.capture(st, include_pressure=include_pressure)
    if mutate is not None:
        mutate(d)
    return V.restore(V.from_bytes(V.to_bytes(d)))


def test_saved_scene_ticks(gpu):
    st = loaded(S.make_scene("612x384", scale_particles=0.5))
    assert st.slot_bound > S.GPU_SPAWN_START  # live slots inside the spawn window
    run_ticks(st, 10, "612x384 loaded")


def test_saved_scene_without_pressure(gpu):
    st = loaded(S.make_scene("612x384", scale_particles=0.25), include_pressure=False)
    assert not st.arrays["air_pv"].any()
    run_ticks(st, 6, "612x384 loaded, no pressure")


def test_saved_scene_with_stacks_overlaps_and_photons(gpu):
    """Every 7th particle duplicated onto its own pixel (stacks), INVIS and FILT
    laid over occupied pixels, photons on matter, fractional positions."""
    rng = np.random.default_rng(11)

    def mutate(d):
        p = d.particles
        dup = p[::7].copy()
        over = p[3::29].copy()
        over["ty"] = np.where(np.arange(len(over)) % 2 == 0, E.pt("INVIS"), E.pt("FILT"))
        phot = p[5::31].copy()
        phot["ty"] = E.pt("PHOT")
        phot["vx"], phot["vy"] = 3.0, 0.0
        phot["life"] = 680
        phot["ctype"] = 0x3FFFFFFF
        p = np.concatenate([p, dup, over, phot])
        p["x"] += rng.uniform(-0.45, 0.45, len(p)).astype(np.float32)
        p["y"] += rng.uniform(-0.45, 0.45, len(p)).astype(np.float32)
        d.particles = p
        d.config["aheat_enable"] = True

    st = loaded(S.make_scene("612x384", scale_particles=0.4), mutate=mutate)
    assert st.arrays["pmap_count"].max() >= 2 and st.arrays["photons"].any()
    run_ticks(st, 12, "612x384 loaded with stacks")


def test_loaded_scene_keeps_state_across_save(gpu):
    """Tick a scene on the GPU, save what the device holds, load it again and
    tick on: the result matches the oracle run over the same saved state."""
    from tpt_hip import lib as L
    st = loaded(S.make_scene("612x384", scale_particles=0.25))
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(5):
            ctx.step(frame)
        ctx.fetch(st)
    again = loaded(st)
    run_ticks(again, 5, "612x384 saved after 5 GPU ticks")
