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


# --- combustion (combustion/fire.rs) and fluid (fluid.rs) passes ---------------

FIRE_T = np.float32(695.15)


def place(parts, pv=0.0, seed=SEED, heat=True):
    """parts: [(x, y, name, {field: value})] in slots 0, 1, ...; pmap rebuilt;
    every air cell at pressure pv."""
    cfg = S.default_config(64, 64, capacity=4096)
    cfg["heat_enable"] = 1 if heat else 0
    st = S.SimState(cfg)
    a = st.arrays
    for k, (x, y, name, fields) in enumerate(parts):
        a["part_ty"][k] = E.pt(name)
        a["part_x"][k], a["part_y"][k] = x, y
        a["part_temp"][k] = fields.get("temp", 300.0)
        for f, v in fields.items():
            if f != "temp":
                a["part_" + f][k] = E.pt(v) if isinstance(v, str) else v
        st.slot_bound = k + 1
    a["air_pv"][:] = pv
    S.rebuild_pmap(st)
    return st


def comb(st, seed=SEED):
    O.OracleState(st).call("oracle_combustion_pass", seed)
    return st.arrays


def fluid(st, seed=SEED):
    O.OracleState(st).call("oracle_fluid_pass", seed)
    return st.arrays


def walled(x, y, skip=()):
    """A ring of METL around (x, y), except the offsets in skip."""
    return [(x + dx, y + dy, "METL", {}) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
            if (dx or dy) and (dx, dy) not in skip]


def test_fire_ignition_gate_needs_space_or_an_explosive():
    """fire.rs:1346-1356: a flame buried in matter cannot light a plain
    flammable; an explosive goes up anyway; one empty neighbour lets it."""
    # OIL (flammable 20) at +2 px from a buried fire; pv 100 saturates the roll
    buried = [(20, 20, "FIRE", {"temp": 700.0})] + walled(20, 20) + [(22, 20, "OIL", {})]
    a = comb(place(buried, pv=100.0))
    assert a["part_ty"][9] == E.pt("OIL")
    buried[-1] = (22, 20, "GUNP", {})
    a = comb(place(buried, pv=60.0))                      # 600 + 60 * 10 >= 1000
    assert a["part_ty"][9] == E.pt("FIRE")
    open_ = [(20, 20, "FIRE", {"temp": 700.0})] + walled(20, 20, skip=[(0, -1)]) + [(22, 20, "OIL", {})]
    a = comb(place(open_, pv=100.0))
    assert a["part_ty"][8] == E.pt("FIRE")


def test_fire_ignition_state_and_pressure():
    """fire.rs:1369-1388 and :340-357: the new fire's temperature is FIRE's
    default plus half the integer flammability, its life 180 + hash % 80,
    tmp and ctype cleared; an explosive ignition adds 0.25 to the burner's
    cell; negative pressure suppresses the roll."""
    st = place([(20, 20, "FIRE", {"temp": 700.0}), (21, 20, "GUNP", {"tmp": 5, "ctype": "WATR"}),
                (22, 20, "OIL", {})], pv=100.0)                           # both rolls saturate
    a = comb(st)
    assert a["part_ty"][1] == E.pt("FIRE") and a["part_ty"][2] == E.pt("FIRE")
    assert a["part_temp"][1] == FIRE_T + np.float32(300.0)
    assert a["part_temp"][2] == FIRE_T + np.float32(10.0)
    for k in (1, 2):
        assert a["part_life"][k] == 180 + hash_u32(k ^ SEED) % 80
        assert a["part_tmp"][k] == 0 and a["part_ctype"][k] == 0
    assert a["air_pv"][(20 // 4) * 16 + 20 // 4] == np.float32(100.25)   # one explosive ignition
    a = comb(place([(20, 20, "FIRE", {"temp": 700.0}), (21, 20, "DUST", {})], pv=-20.0))
    assert a["part_ty"][1] == E.pt("DUST")                                  # 10 - 200 < 0


@pytest.mark.parametrize("seed", [1, 2, 3, 77, 1234])
def test_fire_ignition_roll(seed):
    """fire.rs:333-339: below saturation the roll is chance(seed ^ nid *
    2654435761 ^ n, num, 1000) with n the neighbour's 5x5 index."""
    # every 5x5 neighbour is OIL but the one above, so the flame has space
    parts = [(20, 20, "FIRE", {"temp": 700.0})] + [(20 + dx, 20 + dy, "OIL", {})
                                                    for dy in (-2, -1, 0, 1, 2) for dx in (-2, -1, 0, 1, 2)
                                                    if (dx or dy) and (dx, dy) != (0, -1)]
    a = comb(place(parts, pv=3.0), seed)
    k = 1
    for n in range(25):
        dx, dy = n % 5 - 2, n // 5 - 2
        if not (dx or dy) or (dx, dy) == (0, -1):
            continue
        num = int(np.float32(20.0) + np.float32(3.0) * np.float32(10.0))
        hit = hash_u32(seed ^ ((k * 2654435761) & 0xFFFFFFFF) ^ n) % 1000 < num
        assert (a["part_ty"][k] == E.pt("FIRE")) == hit, (n, k)
        k += 1


def test_fire_heating_composes_in_burner_pixel_order():
    """fire.rs:318-325: each burner hotter than the neighbour moves it 5% of
    the way; burners act in ascending pixel order with their own starting
    temperature."""
    st = place([(20, 20, "FIRE", {"temp": 1000.0}), (22, 20, "FIRE", {"temp": 1000.0}),
                (21, 21, "METL", {"temp": 300.0}), (21, 23, "FIRE", {"temp": 320.0})])
    a = comb(st)
    f = np.float32
    h = f(300.0) + (f(1000.0) - f(300.0)) * f(0.05)        # burner at (20, 20)
    h = h + (f(1000.0) - h) * f(0.05)                        # burner at (22, 20)
    assert a["part_temp"][2] == h                            # the 320 K flame at (21, 23) is colder
    # the two hot burners also heat each other from their starting 1000 K: no change
    assert a["part_temp"][0] == f(1000.0) and a["part_temp"][1] == f(1000.0)
    assert a["part_temp"][3] == f(320.0)                     # three rows below both hot burners


def test_coal_timer_and_wtrv_cooling():
    """fire.rs:188-197: a burner cuts a coal's long timer to 99 on a 1/500
    roll; fire.rs:1011-1013: steam left by burning water sheds 1/1000 of its
    temperature per tick."""
    hits = 0
    for seed in range(150):
        st = place([(20, 20, "FIRE", {"temp": 700.0})] +
                   [(20 + dx, 20 + dy, "COAL", {"life": 150}) for dy in (-2, -1, 1, 2) for dx in (-2, -1, 0, 1, 2)])
        a = comb(st, seed)
        prand = hash_u32(seed ^ ((20 * 64 + 20) * 2654435761 & 0xFFFFFFFF))
        k = 1
        for n in range(25):
            dx, dy = n % 5 - 2, n // 5 - 2
            if dy == 0:
                continue
            nroll = hash_u32(prand ^ ((n * 0x9E3779B9) & 0xFFFFFFFF))
            want = 99 if hash_u32(nroll) % 500 < 1 else 150
            assert a["part_life"][k] == want
            hits += want == 99
            k += 1
    assert hits > 0
    a = comb(place([(10, 10, "WTRV", {"temp": 2000.0, "ctype": "FIRE"}), (30, 30, "WTRV", {"temp": 2000.0})]))
    assert a["part_temp"][0] == np.float32(2000.0) - np.float32(2000.0) / np.float32(1000.0)
    assert a["part_temp"][1] == np.float32(2000.0)


def fluid_roll(seed, idx, salt, n):
    return hash_u32(seed ^ ((idx * salt) & 0xFFFFFFFF) ^ ((n * 0x517CC1B7) & 0xFFFFFFFF))


@pytest.mark.parametrize("seed", range(6))
def test_water_puts_out_fire(seed):
    """fluid.rs:205-216: water kills a neighbouring fire that is not burning
    water (ctype WATR) and dies with it one time in thirty; distilled water
    (:276-283) ignores the ctype."""
    a = fluid(place([(20, 20, "WATR", {}), (21, 20, "FIRE", {}), (19, 20, "FIRE", {"ctype": "WATR"}),
                     (40, 40, "DSTW", {}), (41, 40, "FIRE", {"ctype": "WATR"})]), seed)
    assert a["part_ty"][1] == 0 and a["part_ty"][2] == E.pt("FIRE") and a["part_ty"][4] == 0
    r = fluid_roll(seed, 20 * 64 + 20, 0x2545F491, 5)
    assert (a["part_ty"][0] == 0) == (hash_u32(r ^ 0x2C1B3E4D) % 30 == 0)


def test_two_waters_one_fire_first_conversion_wins():
    """Two waters beside one fire both decide from the fire they saw when the
    pass began; each may die by its own roll."""
    deaths = 0
    for seed in range(60):
        a = fluid(place([(20, 20, "WATR", {}), (21, 20, "FIRE", {}), (22, 20, "WATR", {})]), seed)
        assert a["part_ty"][1] == 0
        for k, (x, n) in enumerate([(20, 5), (22, 3)]):
            r = fluid_roll(seed, 20 * 64 + x, 0x2545F491, n)
            dead = hash_u32(r ^ 0x2C1B3E4D) % 30 == 0
            assert (a["part_ty"][2 * k] == 0) == dead
            deaths += dead
    assert deaths > 0


def test_salt_water_and_ice():
    """fluid.rs:190-198 (WATR + SALT -> SLTW on a 1/50 roll, the salt
    following one try in three) and :357-361 (ICEI with ctype FRZW cools by
    1 K per tick)."""
    conv = 0
    for seed in range(200):
        a = fluid(place([(20, 20, "WATR", {}), (21, 20, "SALT", {})]), seed)
        r = fluid_roll(seed, 20 * 64 + 20, 0x2545F491, 5)
        hit = r % 50 == 0
        assert (a["part_ty"][0] == E.pt("SLTW")) == hit
        assert (a["part_ty"][1] == E.pt("SLTW")) == (hit and hash_u32(r ^ 0x51ED270B) % 3 == 0)
        conv += hit
    assert conv > 0
    a = fluid(place([(20, 20, "ICEI", {"temp": 250.0, "ctype": "FRZW"})]))
    assert a["part_temp"][0] == np.float32(249.0)
