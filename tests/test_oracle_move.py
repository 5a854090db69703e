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
    """A lone WTRV at rest: v = g + diffusion * ((h & 0xFFFF) * 2/65535 - 1)
    with h1 = pcg(seed ^ 0), h2 = pcg(h1) (velocity.rs:160-181), then the
    stochastic unit step fires on each axis iff (h & 0xFFFF)/65535 < |v|
    (velocity.rs:98-111).  Every value is recomputed here in float32."""
    f = np.float32
    lut = E.move_lut()
    t = E.pt("WTRV")
    grav, diff = lut[t * 12 + 9:t * 12 + 11].view(np.float32)
    for seed in (99, 4242, 0x85EBCA6B, 7):
        st = box([("WTRV", 64, 64, 0.0, 0.0)])
        O.OracleState(st).call("oracle_velocity_intent", seed)
        h1 = pcg(seed ^ 0)
        h2 = pcg(h1)
        k = f(2.0) / f(65535.0)
        vx = f(0.0) * lut[t * 12 + 7:t * 12 + 8].view(np.float32)[0] + f(0.0) + f(0.0) + diff * (f(h1 & 0xFFFF) * k - f(1.0))
        vy = f(0.0) + f(0.0) + grav + diff * (f(h2 & 0xFFFF) * k - f(1.0))
        assert st.arrays["part_vx"][0].view(np.uint32) == f(vx).view(np.uint32)
        assert st.arrays["part_vy"][0].view(np.uint32) == f(vy).view(np.uint32)
        unit = f(1.0) / f(65535.0)
        sx = int(np.sign(vx)) if f(h1 & 0xFFFF) * unit < abs(vx) else 0
        sy = int(np.sign(vy)) if f(h2 & 0xFFFF) * unit < abs(vy) else 0
        want = 0xFFFFFFFF if (sx, sy) == (0, 0) else (64 + sy) * 128 + 64 + sx
        assert st.arrays["intent_target"][0] == want, seed


# --- photon refraction at glass (velocity.rs:262-647) -----------------------

def glass_scene(photon, glass_rows=(64, 90), glass_cols=(20, 108)):
    """A 128x128 walled box with a GLAS slab and one photon (x, y, vx, vy, ctype)."""
    c = S.default_config(128, 128, capacity=4096)
    st = S.SimState(c)
    S.wall_ring(st)
    S.derive_block_masks(st)
    a = st.arrays
    k = 0
    for y in range(*glass_rows):
        for x in range(*glass_cols):
            a["part_ty"][k], a["part_x"][k], a["part_y"][k], a["part_temp"][k] = E.pt("GLAS"), x, y, 300.0
            k += 1
    x, y, vx, vy, ctype = photon
    a["part_ty"][k], a["part_x"][k], a["part_y"][k] = E.pt("PHOT"), x, y
    a["part_vx"][k], a["part_vy"][k], a["part_ctype"][k], a["part_temp"][k] = vx, vy, ctype, 300.0
    st.slot_bound = k + 1
    S.rebuild_pmap(st)
    return st, k


# A two-bit wavelength mask (bits 14, 15) is narrow enough that its bin is
# deterministic: lo + hi = 29 (Simulation.cpp:886-922).
NARROW = (1 << 14) | (1 << 15)
N0 = np.float32(1.9) - np.float32(0.07) * (np.float32(29) - np.float32(30)) / np.float32(30)
RATIO_IN = 1.0 / float(N0 * N0)   # glass index as the reference applies it, entering


def run_intent(st):
    o = O.OracleState(st)
    o.call("oracle_velocity_intent", 7)
    return st


def test_refraction_normal_incidence_keeps_velocity():
    st, k = glass_scene((64, 63, 0.0, 1.0, NARROW))
    run_intent(st)
    assert st.arrays["part_vx"][k] == 0.0
    assert st.arrays["part_vy"][k] == pytest.approx(1.0, abs=1e-6)
    assert st.arrays["intent_target"][k] == 64 * 128 + 64          # enters the glass pixel
    assert st.arrays["part_ctype"][k] == NARROW                    # a <5-bit span is kept whole


def test_refraction_entering_glass_obeys_the_index_ratio():
    """Unit speed in: the tangential component scales by the ratio, the
    speed stays 1 (ct2 = 1 - r^2 sin^2), and the photon keeps heading in."""
    st, k = glass_scene((64, 63, 0.6, 0.8, NARROW))
    run_intent(st)
    vx, vy = float(st.arrays["part_vx"][k]), float(st.arrays["part_vy"][k])
    assert vx == pytest.approx(0.6 * RATIO_IN, rel=1e-5)
    assert vx * vx + vy * vy == pytest.approx(1.0, abs=1e-5)
    assert vy > 0.0
    assert st.arrays["intent_target"][k] == 64 * 128 + 65


def test_refraction_total_internal_reflection_mirrors_and_stays():
    """Leaving the glass at 53 degrees from the normal: ratio^2 sin^2 > 1, so
    the velocity is mirrored about the surface and the photon stays put."""
    st, k = glass_scene((64, 64, 0.8, -0.6, NARROW))
    run_intent(st)
    assert float(st.arrays["part_vx"][k]) == pytest.approx(0.8, abs=1e-6)
    assert float(st.arrays["part_vy"][k]) == pytest.approx(0.6, abs=1e-6)
    assert st.arrays["intent_target"][k] == 0xFFFFFFFF


def test_refraction_without_wavelength_kills():
    st, k = glass_scene((64, 63, 0.0, 1.0, 0))
    run_intent(st)
    assert st.arrays["part_ty"][k] == 0


def test_refraction_draws_a_contiguous_band_from_a_wide_mask():
    full = 0x3FFFFFFF
    for seed_slot in range(6):
        st, k = glass_scene((64, 63, 0.3 + 0.05 * seed_slot, 0.9, full))
        run_intent(st)
        got = int(st.arrays["part_ctype"][k])
        assert got != 0 and (got & ~full) == 0
        low = (got & -got).bit_length() - 1
        width = got.bit_length() - low
        assert width in (4, 5) and got == ((1 << width) - 1) << low


# --- WL_DETECT charging and the set_emap closure (mod.rs:348-828) -----------

def wire_scene(cells, particles):
    """128x128 px canvas (32x32 cells): `cells` maps (cx, cy) -> wall id,
    `particles` are (name, x, y)."""
    c = S.default_config(128, 128, capacity=256)
    st = S.SimState(c)
    for (cx, cy), w in cells.items():
        st.arrays["wall_bmap"][cy * 32 + cx] = w
    S.derive_block_masks(st)
    for i, (name, x, y) in enumerate(particles):
        st.arrays["part_ty"][i], st.arrays["part_x"][i], st.arrays["part_y"][i] = E.pt(name), x, y
        st.arrays["part_temp"][i] = 300.0
    st.slot_bound = len(particles)
    S.rebuild_pmap(st)
    return st


def charge(st):
    o = O.OracleState(st)
    o.call("oracle_detect_charge")
    o.call("oracle_emap_closure")
    return st.arrays["wall_emap"].reshape(32, 32)


WL_DETECT, WL_WALLELEC = 3, 1


def test_detector_charges_its_run_only():
    cells = {(x, 10): WL_DETECT for x in range(5, 13)}
    cells.update({(x, 10): WL_DETECT for x in range(14, 17)})   # a separate run after a gap
    em = charge(wire_scene(cells, [("DUST", 7 * 4 + 1, 10 * 4 + 2)]))
    assert (em[10, 5:13] == 16).all()
    assert (em[10, 14:17] == 0).all()
    assert (em & 0xC0).sum() == 0                                # no end markers survive


def test_closure_floods_connected_wire_only():
    cells = {(x, 10): WL_DETECT for x in range(5, 13)}
    cells.update({(12, y): WL_WALLELEC for y in range(11, 15)})  # a one-cell-wide column down
    cells.update({(x, 15): WL_WALLELEC for x in range(12, 19)})  # into another run
    cells.update({(x, 20): WL_WALLELEC for x in range(3, 9)})    # not connected
    em = charge(wire_scene(cells, [("DUST", 6 * 4, 10 * 4)]))
    for (cx, cy) in cells:
        assert em[cy, cx] == (0 if cy == 20 else 16), (cx, cy)
    assert (em & 0xC0).sum() == 0


def test_no_trigger_no_charge():
    cells = {(x, 10): WL_DETECT for x in range(5, 13)}
    em = charge(wire_scene(cells, [("DUST", 40, 80)]))         # particle outside the detector
    assert (em == 0).all()
