"""The element passes (tpt_step_reactions) vs the oracle, bit-exact: the life
pass (fire / plasma / cold flame / TRON burn-in lifetimes, burnout to smoke,
steam or noble gas, sparks settling into every conductor family, the element
kill at zero life, the stasis flag in powered and unpowered WL_STASIS cells),
the combustion pass (FIRE/PLSM heating and ignition of every flammable and
explosive class, the COAL timer, steam cooling and RBDM ignition) and the
fluid pass (water, distilled and salt water, ice and snow against fire,
salt, rock, smoke, plants, freeze-water and rubidium); then whole ticks of
burning scenes."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu
FIELDS = ["part_ty", "part_life", "part_ctype", "part_temp", "part_tmp", "part_flags", "grav_mass_override_set"]
KINDS = ["FIRE", "PLSM", "HFLM", "TRON", "SPRK", "EMBR", "SMKE", "WTRV", "WATR", "METL"]
SPARK_CTYPES = ["WATR", "SLTW", "PSCN", "NSCN", "ETRD", "INWR", "SWCH", "RSST", "METL", "NONE", "DUST"]


def life_scene(seed, workload="612x384"):
    st = S.make_scene(workload, scale_particles=0.5)
    rng = np.random.default_rng(seed)
    a = st.arrays
    live = np.flatnonzero(a["part_ty"][: st.slot_bound])
    pick = rng.choice(live, size=len(live) // 2, replace=False)
    a["part_ty"][pick] = [E.pt(k) for k in rng.choice(KINDS, size=len(pick))]
    a["part_life"][pick] = rng.choice([-1, 0, 1, 2, 5, 200], size=len(pick)).astype(np.int32)
    a["part_tmp"][pick] = rng.integers(0, 8, size=len(pick)).astype(np.int32)
    a["part_temp"][pick] = rng.uniform(300.0, 900.0, size=len(pick)).astype(np.float32)
    sprk = pick[a["part_ty"][pick] == E.pt("SPRK")]
    a["part_ctype"][sprk] = [0 if c == "NONE" else E.pt(c) for c in rng.choice(SPARK_CTYPES, size=len(sprk))]
    plsm = pick[a["part_ty"][pick] == E.pt("PLSM")]
    a["part_ctype"][plsm[::2]] = E.pt("NBLE")
    a["part_ctype"][pick[a["part_ty"][pick] == E.pt("FIRE")]] = 0
    a["part_flags"][pick[::7]] |= 0x20                    # some already in stasis
    cells = rng.choice(a["wall_bmap"].size, size=a["wall_bmap"].size // 6, replace=False)
    a["wall_bmap"][cells] = 18                            # WL_STASIS, half of them powered
    a["wall_emap"][cells] = rng.choice([0, 3, 7, 8, 12], size=len(cells)).astype(np.uint32)
    a["grav_mass_override_set"][rng.choice(a["grav_mass_override_set"].size, 50)] = 1
    return st


CHEM = ["FIRE", "PLSM", "OIL", "DUST", "GUNP", "COAL", "BCOL", "SPNG", "WTRV", "RBDM", "LRBD", "WATR", "DSTW",
        "SLTW", "ICEI", "SNOW", "SALT", "ROCK", "SMKE", "PLNT", "FRZZ", "METL", "NONE"]


def chem_scene(seed, workload="612x384", scale=0.5):
    """Every class the combustion and fluid passes know, packed densely, at
    temperatures around the steam and smoke thresholds, with cells of high,
    low and negative pressure and stasis walls."""
    st = S.make_scene(workload, scale_particles=scale)
    rng = np.random.default_rng(seed)
    a = st.arrays
    live = np.flatnonzero(a["part_ty"][: st.slot_bound])
    names = rng.choice(CHEM, size=len(live))
    a["part_ty"][live] = [E.pt(n) for n in names]
    a["part_temp"][live] = rng.choice([250.0, 300.0, 320.0, 700.0, 1500.0, 2500.0], size=len(live)).astype(np.float32)
    a["part_life"][live] = rng.choice([0, 1, 50, 150, 300], size=len(live)).astype(np.int32)
    a["part_tmp"][live] = rng.integers(0, 4, size=len(live)).astype(np.int32)
    ct = rng.choice(["NONE", "FIRE", "WATR", "FRZW"], size=len(live))
    a["part_ctype"][live] = [0 if c == "NONE" else E.pt(c) for c in ct]
    a["part_vx"][live] = rng.normal(0.0, 1.0, len(live)).astype(np.float32)
    a["air_pv"][:] = rng.choice([-30.0, 0.0, 5.0, 70.0, 120.0], size=a["air_pv"].size).astype(np.float32)
    cells = rng.choice(a["wall_bmap"].size, size=a["wall_bmap"].size // 10, replace=False)
    a["wall_bmap"][cells] = 18
    a["wall_emap"][cells] = rng.choice([0, 9], size=len(cells)).astype(np.uint32)
    S.rebuild_pmap(st)
    return st


@pytest.mark.parametrize("seed", [3, 11, 29])
@pytest.mark.parametrize("heat", [True, False], ids=["heat", "legacy"])
def test_combustion_and_fluid_bit_exact(gpu, seed, heat):
    st = chem_scene(seed)
    st.cfg["heat_enable"] = 1 if heat else 0
    ref = st.copy()
    o = O.OracleState(ref)
    names = FIELDS + ["air_pv"]
    print(f"[combustion+fluid parity] seed {seed} heat {heat}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(3):
            fs = L.frame_seed(frame + seed)
            ctx.step_reactions(fs)
            o.call("oracle_step_reactions", fs)
        ctx.fetch(st, names)
    for n in names:
        a, b = ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)
        assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} differ"


@pytest.mark.parametrize("seed", [1, 7, 2024])
def test_life_pass_bit_exact(gpu, seed):
    st = life_scene(seed)
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[life pass parity] seed {seed}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(3):
            fs = L.frame_seed(frame + seed)
            ctx.step_reactions(fs)
            o.call("oracle_step_reactions", fs)
        ctx.fetch(st, FIELDS)
    for n in FIELDS:
        a, b = ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)
        assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} differ"
    assert not st.arrays["grav_mass_override_set"].any()


@pytest.mark.parametrize("workload,ticks,chem", [("612x384", 12, False), ("2048x2048", 3, False),
                                                 ("612x384", 10, True), ("1600x900", 4, True)])
def test_burning_scene_whole_ticks(gpu, workload, ticks, chem):
    """tpt_step over scenes whose fires spread, burn out and meet water."""
    st = chem_scene(5, workload, 0.5) if chem else life_scene(5, workload)
    ref = st.copy()
    o = O.OracleState(ref)
    names = FIELDS + ["part_x", "part_y", "part_vx", "part_vy", "air_pv", "air_hv", "pmap"]
    print(f"[burning scene] {workload} {ticks} ticks", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            ctx.step(frame)
            o.call("oracle_step", frame)
        ctx.fetch(st, names)
    for n in names:
        a, b = ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)
        assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} differ"
