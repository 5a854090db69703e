"""Oracle: the e == 2 pair pass (reaction_kernel/energy_entry.rs:90-393) and
the drain of the photons it creates (reaction_kernel/shared.rs:516-912).

Pinned by the reference's own known answers where it has them (FILT's
wavelength interaction, rays/mod.rs:227-269; MurmurHash3's published
fmix32(1)), and otherwise by what each arm of Simulation.cpp:1258-1381 fixes
in closed form: a photon absorbed by an empty C5 packs its velocity into the
crystal, H2 rebuilt as ELEC lands on the photon's origin at speed 2, and so
on.  Each scene puts one mover on (64, 63) moving down into an occupant on
(64, 64), which the intent pass sends straight into the occupant's pixel.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S

DEST = 64 * 128 + 64
ORIGIN = 63 * 128 + 64


def pair_scene(mover, occ, mfields=None, ofields=None, cap=20000, extra=()):
    """Occupant in slot 0 at (64, 64), mover in slot 1 at (64, 63), v = (0, 1)."""
    st = S.SimState(S.default_config(128, 128, capacity=cap))
    S.wall_ring(st)
    S.derive_block_masks(st)
    a = st.arrays
    parts = [(occ, 64, 64, ofields or {}), (mover, 64, 63, dict({"part_vy": 1.0}, **(mfields or {})))]
    for k, (name, x, y, f) in enumerate(list(parts) + list(extra)):
        a["part_ty"][k], a["part_x"][k], a["part_y"][k], a["part_temp"][k] = E.pt(name), x, y, 300.0
        for field, v in f.items():
            a[field][k] = v
    st.slot_bound = len(parts) + len(extra)
    S.rebuild_pmap(st)
    return st


def move(st, seed=7):
    O.OracleState(st).call("oracle_step_movement", seed)
    return st.arrays


def test_hash_u32_is_murmur3_fmix32():
    h = O.lib().oracle_hash_u32
    assert h(0) == 0
    assert h(1) == 0x514E28B7          # MurmurHash3 fmix32(1), the published value
    inv1, inv2 = pow(2246822507, -1, 1 << 32), pow(3266489909, -1, 1 << 32)
    rng = np.random.default_rng(0)
    for x in rng.integers(0, 1 << 32, 200, dtype=np.uint64).tolist():
        y = h(x)                        # undo: xorshift-16, *inv2, xorshift-13 (twice), *inv1, xorshift-16
        y ^= y >> 16
        y = (y * inv2) & 0xFFFFFFFF
        y ^= (y >> 13) ^ (y >> 26)
        y = (y * inv1) & 0xFFFFFFFF
        y ^= y >> 16
        assert y == x


def test_interact_wavelengths_reference_cases():
    iw = O.lib().oracle_interact_wavelengths
    # rays/mod.rs:227-269 (test_filt_wavelength_interaction)
    assert iw(0, 0x1234, 300.0, 0x5678) == 0x1234
    assert iw(1, 0x0FF0, 300.0, 0x00FF) == 0x00F0
    assert iw(2, 0x0F00, 300.0, 0x00F0) == 0x0FF0
    assert iw(3, 0x00F0, 300.0, 0x0FF0) == 0x0F00
    assert iw(4, 0, 313.0, 0x1) == 0x2
    assert iw(5, 0, 313.0, 0x2) == 0x1
    assert iw(6, 0x1234, 300.0, 0x5678) == 0x5678
    assert iw(7, 0x0FF0, 300.0, 0x00FF) == 0x0F0F
    assert iw(8, 0, 300.0, 0) == 0x3FFFFFFF
    # Empty FILT ctype: the temperature bin, 0x1F << clamp((T - 273) * 0.025, 0, 25).
    for temp, shift in ((300.0, 0), (353.0, 2), (500.0, 5), (9999.0, 25), (10.0, 0)):
        assert iw(0, 0, temp, 0) == 0x1F << shift
    # Multiply / divide by the lowest set bit of the filter's band.
    assert iw(10, 0x30, 300.0, 0x3) == 0x30
    assert iw(11, 0x30, 300.0, 0x300) == 0x30
    assert iw(11, 0x30, 300.0, 0x7) == 0x0
    assert iw(9, 0x30, 300.0, 0x7) == 0x30          # unknown modes assign the filter's band


def test_photon_through_filt_takes_the_filter_band():
    st = pair_scene("PHOT", "FILT", {"part_ctype": 0x00FF}, {"part_ctype": 0x0FF0, "part_tmp": 1})
    a = move(st)
    assert a["part_ctype"][1] == 0x00F0
    assert (a["part_x"][1], a["part_y"][1]) == (64.0, 64.0)          # shares the filter's pixel
    assert a["part_ctype"][0] == 0x0FF0


def test_photon_into_empty_c5_is_stored():
    st = pair_scene("PHOT", "C5", {"part_ctype": 0xFFC0, "part_vx": 0.2})
    a = move(st)
    assert a["part_ty"][1] == 0                                       # absorbed
    assert a["part_life"][0] == 1 and a["part_ctype"][0] == 0xFFC0
    # tmp packs (vx * 255) low and (vy * 0xFF0000) high; tmp2 the sub-pixel offset (0 here).
    assert a["part_tmp"][0] == (int(np.float32(0.2) * np.float32(255)) & 0xFFFF) | 0xFF0000
    assert a["part_tmp2"][0] == 0
    assert a["photons"][DEST] == 0 and a["photons"][ORIGIN] == 0


def test_photon_into_loaded_c5_collides():
    cvx, cvy = 51, 102                                                # tmp halves, / 255 -> 0.2, 0.4
    st = pair_scene("PHOT", "C5", {"part_ctype": 0xFFC0},
                    {"part_life": 3, "part_ctype": 0x0FC0, "part_tmp": (cvy << 16) | cvx})
    a = move(st)
    assert a["part_ty"][1] == E.pt("PHOT")
    assert a["part_ctype"][1] == (0x0FC0 & 0xFFC0) >> 6
    v = np.hypot(a["part_vx"][1], a["part_vy"][1])
    assert v == pytest.approx(1.0, abs=1e-6)                          # renormalised to the photon's speed
    assert a["part_vx"][1] / a["part_vy"][1] == pytest.approx((cvx / 255) / (1 + cvy / 255), rel=1e-5)
    assert a["part_life"][0] == 0 and a["part_ctype"][0] == 0


@pytest.mark.parametrize("tmp,converts", [(0, True), (1, False)])
def test_photon_splits_h2_into_prot_and_elec(tmp, converts):
    st = pair_scene("PHOT", "H2", {"part_tmp": tmp})
    a = move(st)
    if not converts:
        assert a["part_ty"][1] == E.pt("PHOT") and a["part_ty"][0] == E.pt("H2")
        return
    assert a["part_ty"][1] == E.pt("PROT") and a["part_ctype"][1] == 0 and a["part_tmp2"][1] == 1
    assert (a["part_x"][1], a["part_y"][1]) == (64.0, 64.0)
    assert a["part_ty"][0] == E.pt("ELEC")
    assert (a["part_x"][0], a["part_y"][0]) == (64.0, 63.0)          # the photon's origin
    assert a["part_life"][0] == 680
    d = E.spawn_defaults()[E.pt("ELEC") * 8: E.pt("ELEC") * 8 + 8]
    assert a["part_temp"][0] == np.frombuffer(d[:1].tobytes(), np.float32)[0]
    assert np.hypot(a["part_vx"][0], a["part_vy"][0]) == pytest.approx(2.0, abs=1e-6)


@pytest.mark.parametrize("pv,converts", [(0.0, True), (3.9, True), (4.5, False), (-5.0, False)])
def test_photon_through_invis_becomes_neut_within_its_resistance(pv, converts):
    st = pair_scene("PHOT", "INVIS", {"part_ctype": 0xFF})
    st.arrays["air_pv"][:] = pv
    a = move(st)
    if converts:
        assert a["part_ty"][1] == E.pt("NEUT") and a["part_ctype"][1] == 0
    else:
        assert a["part_ty"][1] == E.pt("PHOT") and a["part_ctype"][1] == 0xFF


@pytest.mark.parametrize("occ,result", [("BIZR", "ELEC"), ("BIZRS", "ELEC")])
def test_photon_into_bizr_becomes_elec(occ, result):
    a = move(pair_scene("PHOT", occ, {"part_ctype": 0xFF}))
    assert a["part_ty"][1] == E.pt(result) and a["part_ctype"][1] == 0


@pytest.mark.parametrize("life,converts", [(0, True), (5, False)])
def test_photon_into_idle_gpmp_becomes_grvt(life, converts):
    a = move(pair_scene("PHOT", "GPMP", {}, {"part_life": life, "part_temp": 500.0}))
    if converts:
        assert a["part_ty"][1] == E.pt("GRVT") and a["part_tmp"][1] == int(np.float32(500.0) - np.float32(273.15))
    else:
        assert a["part_ty"][1] == E.pt("PHOT")


def test_bizr_through_filt_takes_the_filter_band():
    a = move(pair_scene("BIZR", "FILT", {"part_ctype": 0x00FF}, {"part_ctype": 0x0FF0, "part_tmp": 1}))
    assert a["part_ctype"][1] == 0x00F0


@pytest.mark.parametrize("mover,occ", [("NEUT", "GLAS"), ("ELEC", "GLOW"), ("PROT", "INVIS")])
def test_nonzero_rank_nibble_never_passes_the_winner_mask(mover, occ):
    """energy_entry.rs:171-175 compares 27 bits of the winner key, so a mover
    whose weight rank has low bits set (NEUT 2, ELEC 4, PROT 10) never
    resolves a pair: it only shares the pixel."""
    st = pair_scene(mover, occ, {"part_ctype": 0xFF})
    a = move(st)
    assert a["part_ty"][1] == E.pt(mover) and a["part_ty"][0] == E.pt(occ)
    assert np.count_nonzero(a["part_ty"][:20000]) == 2


def zero_rank(st, names):
    """A move LUT whose listed elements have weight rank 0, so that their
    pair arms pass the winner mask (a table the caller may upload)."""
    lut = st.arrays["move_lut"]
    for n in names:
        lut[E.pt(n) * 12 + 2] = 0
    return st


def test_elec_and_prot_arms_with_rank_zero():
    a = move(zero_rank(pair_scene("ELEC", "GLOW"), ["ELEC"]))
    assert a["part_ty"][1] == E.pt("PHOT") and a["part_ctype"][1] == 0x3FFFFFFF
    a = move(zero_rank(pair_scene("PROT", "INVIS"), ["PROT"]))
    assert a["part_ty"][1] == E.pt("NEUT")


def glow_gain_scene(n=100, cap=20000):
    """GLOW on rows 63 and 64 (x = 10 .. 10+n), one photon per column on row 63
    moving down: each crossing has a 1-in-30 chance to fire its laser."""
    extra = []
    for x in range(10, 10 + n):
        extra.append(("GLOW", x, 64, {"part_temp": 300.0 + 20.0 * (x % 7)}))
        extra.append(("GLOW", x, 63, {"part_temp": 280.0 + 40.0 * (x % 5)}))
    for x in range(10, 10 + n):
        extra.append(("PHOT", x, 63, {"part_vy": 1.0, "part_ctype": 0xFF}))
    st = pair_scene("DUST", "DUST", cap=cap, extra=extra)
    a = st.arrays
    a["part_ty"][:2] = 0                                               # the two placeholders
    S.rebuild_pmap(st)
    return st


def test_glow_gain_photons_are_drained_into_the_window_in_requester_order():
    st = glow_gain_scene()
    before = st.copy()
    a = move(st)
    b = before.arrays
    glow64 = [k for k in range(st.slot_bound) if b["part_ty"][k] == E.pt("GLOW") and b["part_y"][k] == 64]
    fired = [k for k in glow64 if a["part_life"][k] == 120]
    assert 0 < len(fired) < 20
    created = np.flatnonzero(a["part_ty"][1024:17408]) + 1024
    assert list(created) == list(range(1024, 1024 + len(fired)))      # the lowest free slots, in order
    assert st.slot_bound == 17408
    # Requesters are photons in ascending slot order, i.e. ascending x here.
    xs = [int(a["part_x"][s]) for s in created]
    assert xs == sorted(xs) and len(set(xs)) == len(xs)
    for s in created:
        x, y = int(a["part_x"][s]), int(a["part_y"][s])
        assert y == 63 and (x, 64) in {(int(b["part_x"][k]), 64) for k in fired}
        gain = [k for k in range(st.slot_bound) if b["part_ty"][k] == E.pt("GLOW")
                and (b["part_x"][k], b["part_y"][k]) == (x, 63)][0]
        t = b["part_temp"][gain]
        assert a["part_temp"][s] == t
        assert a["part_ctype"][s] == 0x1F << min(max(int((t - np.float32(273.0)) * np.float32(0.25)), 0), 25)
        assert (a["part_vx"][s], a["part_vy"][s]) == (0.0, 1.0)
        assert a["part_life"][s] == 680 and a["part_flags"][s] == 0     # PHOT's DefaultProperties
        assert a["photons"][y * 128 + x] != 0                          # scattered by the rebuild


def test_drain_skips_taken_window_slots():
    st = glow_gain_scene()
    st.arrays["part_ty"][1024] = E.pt("METL")                          # not free
    st.arrays["part_x"][1024], st.arrays["part_y"][1024] = 5, 5
    st.arrays["part_flags"][1025] = 0x04                               # portal-held: not free either
    st.slot_bound = 1026
    S.rebuild_pmap(st)
    a = move(st)
    created = np.flatnonzero(a["part_ty"][1026:17408]) + 1026
    assert len(created) > 0 and created[0] == 1026
    assert a["part_ty"][1025] == 0


def test_cherenkov_photon_with_rank_zero_neutrons():
    """NEUT leaving glass into glass, fast enough (|v| >= 1.44): one in ten
    emits a photon at the neutron's origin, at speed 1.269, wavelength 0xF80."""
    extra = [("GLAS", x, y, {"part_temp": 350.0}) for y in range(61, 68) for x in range(10, 110)]
    extra += [("NEUT", x, 63, {"part_vy": 1.5}) for x in range(10, 110)]
    st = zero_rank(pair_scene("DUST", "DUST", extra=extra), ["NEUT"])
    st.arrays["part_ty"][:2] = 0
    S.rebuild_pmap(st)
    a = move(st)
    created = np.flatnonzero(a["part_ty"][1024:17408]) + 1024
    assert 0 < len(created) < 40
    for s in created:
        assert a["part_ty"][s] == E.pt("PHOT") and a["part_ctype"][s] == 0xF80
        assert a["part_y"][s] == 63 and a["part_temp"][s] == np.float32(350.0)
        assert np.hypot(a["part_vx"][s], a["part_vy"][s]) == pytest.approx(1.269, rel=1e-6)
        # v' = (vx, vy) -/+ 2.5 * (vy, -vx), renormalised: it leaves at
        # atan(2.5) to the neutron's path (here nearly straight down).
        assert abs(a["part_vx"][s] / a["part_vy"][s]) == pytest.approx(2.5, rel=2e-2)
