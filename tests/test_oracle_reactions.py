"""Pins the oracle's life pass (oracle/reactions.c) against the reference's
own unit tests of life.rs (fire_burnout_transitions,
plsm_burnout_transitions, default_life_ranges; reaction_kernel/life.rs:207-270)
and closed forms written out from life.rs:47-205, with hash_u32 (MurmurHash3's
finaliser, reaction_kernel/shared.rs:1015-1023) recomputed here."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S

SEED = 0x1234ABCD


def hash_u32(x):
    h = x & 0xFFFFFFFF
    h ^= h >> 16
    h = (h * 2246822507) & 0xFFFFFFFF
    h ^= h >> 13
    h = (h * 3266489909) & 0xFFFFFFFF
    h ^= h >> 16
    return h


def run(parts, seed=SEED, walls=None):
    """parts: [(slot, name, life, tmp, temp, ctype)] at (20 + slot, 20);
    returns the state after the life pass."""
    st = S.SimState(S.default_config(256, 64, capacity=4096))
    a = st.arrays
    for (k, name, life, tmp, temp, ctype) in parts:
        a["part_ty"][k] = E.pt(name) if isinstance(name, str) else name
        a["part_x"][k], a["part_y"][k] = 20 + k % 200, 20 + (k // 200) % 40
        a["part_life"][k], a["part_tmp"][k], a["part_temp"][k] = life, tmp, temp
        a["part_ctype"][k] = E.pt(ctype) if isinstance(ctype, str) else ctype
        st.slot_bound = max(st.slot_bound, k + 1)
    for (cell, w, e) in walls or []:
        a["wall_bmap"][cell], a["wall_emap"][cell] = w, e
    O.OracleState(st).call("oracle_life_pass", seed)
    return a


def test_fire_burnout_transitions():
    """H2+O2-fed fire leaves steam; cold fire leaves smoke with a 250..269
    life; hot unmixed fire burns out completely."""
    a = run([(0, "FIRE", 1, 3, 700.0, 0), (1, "FIRE", 1, 0, 800.0, 0)] +
            [(k, "FIRE", 1, 0, 500.0, 0) for k in range(2, 66)])
    assert (a["part_ty"][0], a["part_life"][0], a["part_ctype"][0]) == (E.pt("WTRV"), 0, E.pt("FIRE"))
    assert a["part_ty"][1] == 0
    for k in range(2, 66):
        assert (a["part_ty"][k], a["part_ctype"][k], a["part_tmp"][k]) == (E.pt("SMKE"), 0, 0)
        assert 250 <= a["part_life"][k] <= 269
        assert a["part_life"][k] == 250 + hash_u32(k ^ SEED ^ 0x5ACE123) % 20


def test_plsm_burnout_transitions():
    a = run([(0, "PLSM", 1, 0, 9999.0, "NBLE"), (1, "PLSM", 1, 3, 9999.0, 0), (2, "PLSM", 1, 0, 9999.0, 0),
             (3, "PLSM", 5, 0, 9999.0, 0)])
    assert (a["part_ty"][0], a["part_life"][0], a["part_ctype"][0]) == (E.pt("NBLE"), 0, 0)
    assert (a["part_ty"][1], a["part_life"][1], a["part_ctype"][1]) == (E.pt("WTRV"), 0, E.pt("FIRE"))
    assert a["part_ty"][2] == 0
    assert (a["part_ty"][3], a["part_life"][3]) == (E.pt("PLSM"), 5)       # still burning


@pytest.mark.parametrize("seed", [0, 42, 100, 999])
def test_default_life_ranges(seed):
    """A fire, plasma, cold flame or TRON created with life 0 starts a fresh
    burn of 180..259, 50..199 or 25..49 ticks."""
    a = run([(k, name, 0, 0, 400.0, 0) for k, name in enumerate(["FIRE", "PLSM", "HFLM", "TRON"] * 8)], seed=seed)
    for k, name in enumerate(["FIRE", "PLSM", "HFLM", "TRON"] * 8):
        lo, span, salt = {"FIRE": (180, 80, 0), "PLSM": (50, 150, 0xA5A5), "HFLM": (50, 150, 0xA5A5),
                          "TRON": (25, 25, 0x3C3C)}[name]
        assert a["part_ty"][k] == E.pt(name)
        assert a["part_life"][k] == lo + hash_u32(k ^ seed ^ salt) % span
        assert lo <= a["part_life"][k] < lo + span


def test_spark_settles_into_its_conductor():
    """SPRK.cpp:62-88 as life.rs:144-180 ports it: a spent spark returns to
    its ctype (METL when none) with the refractory life, the water/logic
    family at room temperature; RSST vanishes."""
    a = run([(0, "SPRK", 0, 0, 500.0, "WATR"), (1, "SPRK", 0, 0, 500.0, 0), (2, "SPRK", -1, 0, 500.0, "SLTW"),
             (3, "SPRK", 0, 0, 500.0, "SWCH"), (4, "SPRK", 0, 0, 500.0, "RSST"), (5, "SPRK", 0, 0, 500.0, "PSCN"),
             (6, "SPRK", 2, 0, 500.0, "METL")])
    want = [("WATR", 64, 293.15), ("METL", 4, 500.0), ("SLTW", 54, 293.15), ("SWCH", 14, 500.0), (None, 0, 0),
            ("PSCN", 4, 293.15)]
    for k, (name, life, temp) in enumerate(want):
        if name is None:
            assert a["part_ty"][k] == 0
            continue
        assert (a["part_ty"][k], a["part_life"][k], a["part_ctype"][k]) == (E.pt(name), life, 0), k
        assert a["part_temp"][k] == np.float32(temp), k
    assert (a["part_ty"][6], a["part_life"][6]) == (E.pt("SPRK"), 2)      # still sparking


def test_life_kill_and_stasis():
    """An element with PROP_LIFE_KILL and no life left dies (EMBR); a particle
    in an unpowered WL_STASIS cell only gets the stasis flag, and loses it in
    a powered one."""
    assert E.element_props_table()[E.pt("EMBR") * 32 + 1] & 0x8000
    a = run([(80, "EMBR", 0, 0, 400.0, 0), (81, "EMBR", 3, 0, 400.0, 0), (2, "FIRE", 1, 0, 500.0, 0),
             (40, "FIRE", 1, 0, 500.0, 0)],
            walls=[((20 // 4) * 64 + (22 // 4), 18, 0), ((20 // 4) * 64 + (60 // 4), 18, 9)])
    assert a["part_ty"][80] == 0 and a["part_ty"][81] == E.pt("EMBR")
    assert a["part_ty"][2] == E.pt("FIRE") and a["part_flags"][2] & 0x20 and a["part_life"][2] == 1
    assert a["part_ty"][40] == E.pt("SMKE") and not a["part_flags"][40] & 0x20
