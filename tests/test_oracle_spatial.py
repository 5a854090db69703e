"""Pins the RecalcFreeParticles/CheckStacking oracle (oracle/spatial.c).

Invariants the map rebuild must satisfy (buffers/spatial.rs:81-275):
every occupied pmap pixel names a live matter particle standing on it;
pmap_count counts the matter on each pixel (minus the THDR/EMBR/FIGH/PLSM
exclusions); the highest slot owns a shared pixel; INVIS only takes an empty
pixel; PROP_LIFE_DEC decrements life and PROP_LIFE_KILL kills at zero;
CheckStacking turns a stack into one NBHL with tmp = stack size."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S


def test_hash_known_answers():
    # lowbias32 with the 0x9E3779B9 pre-mix: computed here independently.
    def h(x):
        x = (x ^ 0x9E3779B9) & 0xFFFFFFFF
        x ^= x >> 16
        x = (x * 0x7FEB352D) & 0xFFFFFFFF
        x ^= x >> 15
        x = (x * 0x846CA68B) & 0xFFFFFFFF
        return x ^ (x >> 16)
    for v in (0, 1, 0xDEADBEEF, 0x9E3779B9, 12345678):
        assert O.lib().oracle_lowbias_hash(v) == h(v)


def test_rebuild_invariants_on_scene():
    st = S.make_scene("612x384", scale_particles=0.5)
    st.arrays["pmap"][:] = 12345  # stale content must be cleared
    O.OracleState(st).call("oracle_recalc_free_particles", 1, 0)
    pm, cnt = st.arrays["pmap"], st.arrays["pmap_count"]
    occ = np.flatnonzero(pm)
    slots = (pm[occ] >> 9) - 1
    assert np.all(st.arrays["part_ty"][slots] == (pm[occ] & 0x1FF))
    xs = st.arrays["part_x"][slots].astype(np.int64)
    ys = st.arrays["part_y"][slots].astype(np.int64)
    assert np.array_equal(ys * st.cfg["xres"] + xs, occ)
    live = np.flatnonzero(st.arrays["part_ty"][: st.slot_bound])
    assert len(occ) == len(live) and cnt.sum() == len(live) and cnt.max() == 1


def _state_with(parts, xres=64, yres=64):
    st = S.SimState(S.default_config(xres, yres, capacity=64))
    for i, (name, x, y, life) in enumerate(parts):
        st.arrays["part_ty"][i] = E.pt(name)
        st.arrays["part_x"][i] = x
        st.arrays["part_y"][i] = y
        st.arrays["part_life"][i] = life
    st.slot_bound = len(parts)
    return st


def test_shared_pixel_highest_slot_and_invis():
    st = _state_with([("DUST", 10, 10, 0), ("WATR", 10, 10, 0), ("INVIS", 10, 10, 0),
                      ("INVIS", 20, 20, 0), ("INVIS", 20, 20, 0), ("PHOT", 30, 30, 0), ("PHOT", 30, 30, 0)])
    O.OracleState(st).call("oracle_recalc_free_particles", 0, 0)
    pm = st.arrays["pmap"].reshape(64, 64)
    assert pm[10, 10] >> 9 == 2            # WATR (slot 1) last non-INVIS writer
    assert pm[20, 20] >> 9 == 4            # lowest INVIS takes the empty pixel
    assert st.arrays["photons"].reshape(64, 64)[30, 30] >> 9 == 7
    assert st.arrays["pmap_count"].reshape(64, 64)[10, 10] == 3


def test_life_decay_and_kill():
    # FIRE: PROP_LIFE_DEC | PROP_LIFE_KILL
    st = _state_with([("FIRE", 10, 10, 5), ("FIRE", 12, 10, 1), ("FIRE", 14, 10, 0), ("DUST", 16, 10, 0)])
    O.OracleState(st).call("oracle_recalc_free_particles", 0, 0)
    assert list(st.arrays["part_life"][:4]) == [4, 0, 0, 0]
    assert list(st.arrays["part_ty"][:4]) == [E.pt("FIRE"), 0, 0, E.pt("DUST")]
    assert st.arrays["pmap"].reshape(64, 64)[10, 12] == 0


def test_stacking_collapses_to_nbhl():
    st = _state_with([("DUST", 10, 10, 0)] * 40)
    O.OracleState(st).call("oracle_recalc_free_particles", 7, 1)
    ty = st.arrays["part_ty"][:40]
    assert np.count_nonzero(ty) == 1 and ty[39] == E.pt("NBHL")   # pmap owner = highest slot
    assert st.arrays["part_tmp"][39] == 40 and st.arrays["part_temp"][39] == np.float32(9999.0)
    assert st.arrays["pmap_count"].reshape(64, 64)[10, 10] == 64 * 64
