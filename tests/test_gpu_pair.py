"""HIP e == 2 pair pass and the drain of its photons vs the oracle, bit for
bit: every arm of energy_entry.rs on its own scene, GLOW gain and Cherenkov
creations drained into the device window, and a mixed swarm through whole
ticks (with the reference's weight ranks and with ranks zeroed so that the
NEUT/ELEC/PROT arms run too)."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import lib as L
from tpt_hip import scene as S
from test_gpu_move import FLOAT_FIELDS, INT_FIELDS, compare
from test_oracle_pair import glow_gain_scene, pair_scene, zero_rank

pytestmark = pytest.mark.gpu
PAIR_FIELDS = INT_FIELDS + FLOAT_FIELDS + ["part_tmp2", "part_tmp3", "part_tmp4", "part_dcolour", "wall_emap"]


def run_both(st, ticks, whole=False, label=""):
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[pair parity] {label} ticks={ticks} whole={whole}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            if whole:
                ctx.step(frame)
                o.call("oracle_step", frame)
            else:
                seed = L.frame_seed(frame)
                ctx.recalc_free_particles(seed, False)
                o.call("oracle_recalc_free_particles", seed, 0)
                ctx.step_movement(seed)
                o.call("oracle_step_movement", seed)
        ctx.fetch(st)
    compare(ref, st, PAIR_FIELDS, label)
    assert st.slot_bound == ref.slot_bound
    return ref


ARMS = [("PHOT", "FILT", {"part_ctype": 0x00FF}, {"part_ctype": 0x0FF0, "part_tmp": 1}),
        ("PHOT", "C5", {"part_ctype": 0xFFC0, "part_vx": 0.2}, {}),
        ("PHOT", "C5", {"part_ctype": 0xFFC0}, {"part_life": 3, "part_ctype": 0x0FC0, "part_tmp": (102 << 16) | 51}),
        ("PHOT", "H2", {"part_tmp": 0}, {}),
        ("PHOT", "INVIS", {"part_ctype": 0xFF}, {}),
        ("PHOT", "BIZR", {"part_ctype": 0xFF}, {}),
        ("PHOT", "GPMP", {}, {"part_temp": 500.0}),
        ("BIZR", "FILT", {"part_ctype": 0x00FF}, {"part_ctype": 0x0FF0, "part_tmp": 7}),
        ("NEUT", "GLAS", {}, {}),
        ("ELEC", "GLOW", {}, {})]


@pytest.mark.parametrize("arm", range(len(ARMS)))
def test_each_arm_bit_exact(gpu, arm):
    mover, occ, mf, of = ARMS[arm]
    run_both(pair_scene(mover, occ, mf, of), 3, label=f"{mover}+{occ}")


def test_glow_gain_drain_bit_exact(gpu):
    ref = run_both(glow_gain_scene(), 2, label="GLOW gain")
    assert np.count_nonzero(ref.arrays["part_ty"][1024:17408]) > 0


def test_cherenkov_with_rank_zero_bit_exact(gpu):
    extra = [("GLAS", x, y, {"part_temp": 350.0}) for y in range(61, 68) for x in range(10, 110)]
    extra += [("NEUT", x, 63, {"part_vy": 1.5}) for x in range(10, 110)]
    st = zero_rank(pair_scene("DUST", "DUST", extra=extra), ["NEUT"])
    st.arrays["part_ty"][:2] = 0
    S.rebuild_pmap(st)
    ref = run_both(st, 2, label="Cherenkov")
    assert np.count_nonzero(ref.arrays["part_ty"][1024:17408]) > 0


def swarm(seed, ranks_zero):
    """612x384 workload (30%) with blocks of every pair occupant and a swarm
    of photons, neutrons, electrons and protons."""
    st = S.make_scene("612x384", scale_particles=0.3, newton_grav=False)
    rng = np.random.default_rng(seed)
    a = st.arrays
    occupied = set(zip(a["part_x"][: st.slot_bound].astype(int), a["part_y"][: st.slot_bound].astype(int)))
    add = []
    for name in ["FILT", "C5", "H2", "INVIS", "BIZR", "BIZRS", "GPMP", "GLOW", "GLAS", "BGLA"] * 2:
        x0, y0 = int(rng.integers(20, 560)), int(rng.integers(20, 330))
        for y in range(y0, y0 + int(rng.integers(4, 16))):
            for x in range(x0, x0 + int(rng.integers(4, 30))):
                if (x, y) not in occupied:
                    occupied.add((x, y))
                    f = {"part_tmp": int(rng.integers(0, 12)), "part_temp": float(rng.uniform(280, 900)),
                         "part_ctype": int(rng.choice([0, 0x3FC0, 0xFF]))}
                    if name == "C5":
                        f["part_life"] = int(rng.integers(0, 2))
                    if name == "GLOW":
                        f["part_life"] = 0
                    add.append((name, x, y, f))
    for _ in range(4000):
        name = str(rng.choice(["PHOT", "PHOT", "NEUT", "ELEC", "PROT"]))
        th = rng.uniform(0, 2 * np.pi)
        sp = 3.0 if name == "PHOT" else float(rng.uniform(0.5, 2.5))
        add.append((name, int(rng.integers(10, 600)), int(rng.integers(10, 370)),
                    {"part_vx": sp * np.cos(th), "part_vy": sp * np.sin(th), "part_life": 680,
                     "part_ctype": int(rng.choice([0x3FFFFFFF, 0xFFC0, 0x3F])), "part_tmp": int(rng.integers(0, 2)),
                     "part_temp": 300.0}))
    base = st.slot_bound
    for j, (name, x, y, f) in enumerate(add):
        s = base + j
        a["part_ty"][s], a["part_x"][s], a["part_y"][s] = E.pt(name), x, y
        for k, v in f.items():
            a[k][s] = v
    st.slot_bound = base + len(add)
    S.rebuild_pmap(st)
    if ranks_zero:
        zero_rank(st, ["NEUT", "ELEC", "PROT"])
    return st


@pytest.mark.parametrize("ranks_zero", [False, True])
def test_swarm_movement_bit_exact(gpu, ranks_zero):
    run_both(swarm(61, ranks_zero), 6, label=f"swarm ranks_zero={ranks_zero}")


def test_swarm_whole_ticks_bit_exact(gpu):
    run_both(swarm(67, True), 8, whole=True, label="swarm whole ticks")
