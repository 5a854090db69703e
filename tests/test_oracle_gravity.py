"""Pins the gravity oracle (oracle/gravity.c).

* The reference's own FFT tests, restated: gravity_gpu/fft.rs:357-393
  (radix-2 passes vs a naive 2-D DFT on 8x8, tolerance 1e-3) and
  fft.rs:394-412 (forward then inverse scales by the sample count, 4x8).
* The FP64 solve against an O(C^2) brute-force convolution
  sum_m mass[m] * (-M_GRAV) * d/|d|^3 (the device-vs-brute-force check
  described at gravity_gpu/mod.rs:38-41) and the analytic point-mass field.
* The float32 restatement of the reference arithmetic against the FP64 solve.
* Mask flood fill (gravity_kernel.rs:658-710), fixed-point mass deposit
  (gravity_kernel.rs:336-345) and Loss + force (gravity_kernel.rs:434-507).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S

G = float(np.float32(6.673e-1))


def naive_dft2(z, sign):
    h, w = z.shape
    ky, y = np.meshgrid(np.arange(h), np.arange(h), indexing="ij")
    kx, x = np.meshgrid(np.arange(w), np.arange(w), indexing="ij")
    Fy = np.exp(sign * 2j * np.pi * ky * y / h)
    Fx = np.exp(sign * 2j * np.pi * kx * x / w)
    return Fy @ z @ Fx.T


def test_reference_fft_matches_naive_dft_8x8():
    w = h = 8
    x, y = np.meshgrid(np.arange(w), np.arange(h))
    re = ((x * 7 + y * 3) % 11).astype(np.float32) - 5.0
    im = ((x * 5 + y * 9) % 7).astype(np.float32) - 3.0
    data = np.empty(w * h * 2, np.float32)
    data[0::2], data[1::2] = re.ravel(), im.ravel()
    O.fft_f32(data, h, w, w, 1, 1.0, 3)
    O.fft_f32(data, w, h, 1, w, 1.0, 3)
    want = naive_dft2(re + 1j * im, -1.0).ravel()
    assert np.max(np.abs(data[0::2] - want.real)) < 1e-3
    assert np.max(np.abs(data[1::2] - want.imag)) < 1e-3


def test_reference_fft_forward_inverse_scales_by_count():
    w, h = 4, 8
    data = np.zeros(w * h * 2, np.float32)
    data[0::2] = np.arange(w * h, dtype=np.float32) * 0.25 - 1.0
    orig = data.copy()
    for sign in (1.0, -1.0):
        O.fft_f32(data, h, w, w, 1, sign, 3)
        O.fft_f32(data, w, h, 1, w, sign, 3)
    assert np.max(np.abs(data[0::2] / (w * h) - orig[0::2])) < 1e-3


def small_state(xres=96, yres=80, seed=3, mask_box=False):
    st = S.SimState(S.default_config(xres, yres, capacity=1024))
    rng = np.random.default_rng(seed)
    n = st.xcells * st.ycells
    st.arrays["grav_mask"][:] = 0xFFFFFFFF
    m = np.zeros(n, np.int32)
    idx = rng.choice(n, size=12, replace=False)
    m[idx] = (rng.uniform(-50, 100, size=12) * 65536).astype(np.int32)
    st.arrays["grav_mass_prev"][:] = m
    if mask_box:
        st.arrays["grav_mask"].reshape(st.ycells, st.xcells)[5:9, 5:12] = 0
    return st


def brute_force(st):
    xc, yc = st.xcells, st.ycells
    mask = st.arrays["grav_mask"].reshape(yc, xc) != 0
    m = st.arrays["grav_mass_prev"].reshape(yc, xc).astype(np.float64) / 65536.0 * mask
    ys, xs = np.nonzero(m)
    Y, X = np.mgrid[0:yc, 0:xc]
    fx = np.zeros((yc, xc))
    fy = np.zeros((yc, xc))
    for y0, x0 in zip(ys, xs):
        dx, dy = X - x0, Y - y0
        r3 = np.power(dx * dx + dy * dy, 1.5)
        r3[y0, x0] = np.inf
        fx += m[y0, x0] * (-G) * dx / r3
        fy += m[y0, x0] * (-G) * dy / r3
    return (fx * mask).ravel(), (fy * mask).ravel()


@pytest.mark.parametrize("mask_box", [False, True])
def test_fp64_solve_equals_brute_force_convolution(mask_box):
    st = small_state(mask_box=mask_box)
    fx, fy = O.gravity_solve_f64(st)
    bx, by = brute_force(st)
    scale = max(np.abs(bx).max(), np.abs(by).max())
    assert np.max(np.abs(fx - bx)) / scale < 1e-9
    assert np.max(np.abs(fy - by)) / scale < 1e-9


def test_point_mass_analytic_field():
    st = S.SimState(S.default_config(64, 64, capacity=1024))
    st.arrays["grav_mask"][:] = 0xFFFFFFFF
    st.arrays["grav_mass_prev"].reshape(16, 16)[7, 9] = 3 * 65536
    fx, fy = O.gravity_solve_f64(st)
    fx, fy = fx.reshape(16, 16), fy.reshape(16, 16)
    # Two cells to the right: attraction back toward the mass, -G*M/r^2 along -x.
    assert fx[7, 11] == pytest.approx(-G * 3 * 2 / 8.0, rel=1e-9)
    assert fy[7, 11] == pytest.approx(0.0, abs=1e-12)
    assert fy[4, 9] == pytest.approx(-G * 3 * (-3) / 27.0, rel=1e-9)
    assert abs(fx[7, 9]) < 1e-12 and abs(fy[7, 9]) < 1e-12


def test_f32_reference_arithmetic_within_1e4_of_fp64():
    st = small_state(xres=612, yres=384)
    O.OracleState(st).call("oracle_gravity_solve")
    fx, fy = O.gravity_solve_f64(st)
    scale = max(np.abs(fx).max(), np.abs(fy).max())
    assert np.max(np.abs(st.arrays["grav_fx"] - fx)) / scale < 1e-4
    assert np.max(np.abs(st.arrays["grav_fy"] - fy)) / scale < 1e-4


def test_gravity_mask_flood_fill():
    xc, yc = 20, 16
    wall = np.zeros(xc * yc, np.uint32)
    w2 = wall.reshape(yc, xc)
    w2[4, 4:10] = 14
    w2[9, 4:10] = 14
    w2[4:10, 4] = 14
    w2[4:10, 9] = 14
    mask = O.gravity_mask(wall, xc, yc).reshape(yc, xc)
    assert not mask[5:9, 5:9].any()             # sealed box interior
    assert not mask[w2 == 14].any()             # the walls themselves
    assert (mask[w2 != 14][:10] == 0xFFFFFFFF).all() and mask[0, 0] == 0xFFFFFFFF
    assert np.count_nonzero(mask) == xc * yc - 20 - 16  # 6x6 ring of 20 walls, 4x4 inside


def test_mass_deposit_fixed_point():
    st = S.SimState(S.default_config(64, 64, capacity=64))
    st.arrays["grav_mask"][:] = 0xFFFFFFFF
    a = st.arrays
    a["part_ty"][0], a["part_x"][0], a["part_y"][0], a["part_tmp"][0] = E.pt("GRVT"), 10, 10, 50
    a["part_ty"][1], a["part_x"][1], a["part_y"][1], a["part_tmp"][1] = E.pt("NWHL"), 30, 30, 0
    a["part_ty"][2], a["part_x"][2], a["part_y"][2] = E.pt("DUST"), 11, 9
    st.slot_bound = 3
    O.OracleState(st).call("oracle_build_gravity_mass")
    m = a["grav_mass"].reshape(16, 16)
    assert m[2, 2] == 10 * 65536               # 0.2 * clamp(50) = 10
    assert m[7, 7] == -int(np.float32(0.1) * np.float32(65536) + np.float32(0.5))  # hole default 0.1
    assert np.count_nonzero(m) == 2            # DUST deposits nothing in reference mode
    st.cfg["grav_mass_mode"] = 1
    a["grav_mass"][:] = 0
    O.OracleState(st).call("oracle_build_gravity_mass")
    assert m[2, 2] == (E.ELEMENTS[E.pt("GRVT")]["weight"] + 85) * 65536


def test_apply_loss_then_force():
    st = S.SimState(S.default_config(64, 64, capacity=64))
    a = st.arrays
    for i, (name, v) in enumerate((("DUST", 1.0), ("METL", 1.0))):
        a["part_ty"][i], a["part_x"][i], a["part_y"][i] = E.pt(name), 20, 20
        a["part_vx"][i] = a["part_vy"][i] = v
    a["part_ty"][2], a["part_x"][2], a["part_y"][2], a["part_vx"][2] = E.pt("DUST"), 2, 20, 1.0  # margin
    st.slot_bound = 3
    a["grav_fx"].reshape(16, 16)[5, 5] = 0.5
    a["grav_fy"].reshape(16, 16)[5, 5] = -0.25
    O.OracleState(st).call("oracle_apply_gravity")
    assert a["part_vx"][0] == pytest.approx(0.8 + 0.5) and a["part_vy"][0] == pytest.approx(0.8 - 0.25)
    assert a["part_vx"][1] == 0.0                # METL: Loss 0, solids get no force
    assert a["part_vx"][2] == 1.0                # inside the 4-px margin: untouched


# --- PGRV/NGRV tool wells and element mass overrides (gravity_kernel.rs:241-262, 379-391) ---

def well_state(prev=None, tool=None, particles=()):
    st = S.SimState(S.default_config(128, 128, capacity=64))
    st.cfg["newton_grav"] = 1
    st.arrays["grav_mask"][:] = 0xFFFFFFFF
    for cell, v in (prev or {}).items():
        st.arrays["grav_mass_prev"][cell] = v
    for cell, v in (tool or {}).items():
        st.arrays["grav_tool_mass"][cell] = v
    for i, (name, x, y, tmp) in enumerate(particles):
        a = st.arrays
        a["part_ty"][i], a["part_x"][i], a["part_y"][i], a["part_tmp"][i] = E.pt(name), x, y, tmp
        st.slot_bound = i + 1
    return st


def field_of(masses, n=32):
    """FP64 field of {cell: fixed-point mass} on an all-open 32x32-cell grid."""
    st = S.SimState(S.default_config(4 * n, 4 * n, capacity=64))
    st.arrays["grav_mask"][:] = 0xFFFFFFFF
    for cell, v in masses.items():
        st.arrays["grav_mass_prev"][cell] = v
    return O.gravity_solve_f64(st)


def close_field(st, want, rel):
    fx, fy = want
    scale = max(np.abs(fx).max(), np.abs(fy).max())
    return (np.abs(st.arrays["grav_fx"] - fx).max() / scale < rel and
            np.abs(st.arrays["grav_fy"] - fy).max() / scale < rel)


@pytest.mark.parametrize("units,q", [(2.5, 3), (-2.5, -3), (1.49, 1), (-1.49, -1), (0.5, 1), (7.0, 7)])
def test_tool_well_rounds_half_away_into_the_solve(units, q):
    """A well of `units`/65536 mass enters the solve as round-half-away(units)
    fixed-point units (tool * 65536 +- 0.5, truncated); the field is linear in
    that integer, so 3 vs 2 or 1 vs 2 units differ by a third or more."""
    c = 12 * 32 + 9
    st = well_state(tool={c: np.float32(units) / np.float32(65536.0)})
    O.OracleState(st).call("oracle_step_gravity")
    assert close_field(st, field_of({c: q}), 1e-3)
    assert not close_field(st, field_of({c: q + (1 if q > 0 else -1)}), 1e-2)
    assert not st.arrays["grav_tool_mass"].any()           # spent


def test_tool_well_replaces_its_cell_and_keeps_the_others():
    """The seed writes only non-zero tool cells: other snapshot cells keep the
    previous tick's particle mass, and a well on a massive cell replaces it."""
    a, b = 10 * 32 + 10, 20 * 32 + 21
    st = well_state(prev={a: 40 * 65536, b: 25 * 65536}, tool={b: 3.0})
    O.OracleState(st).call("oracle_step_gravity")
    assert close_field(st, field_of({a: 40 * 65536, b: 3 * 65536}), 1e-4)
    assert not close_field(st, field_of({a: 40 * 65536, b: 28 * 65536}), 1e-3)


def test_override_replaces_the_built_mass_and_joins_the_snapshot():
    """apply_grav_mass_override_kernel: a flagged cell takes the assigned
    value over the particle deposit (GRVT tmp 50 -> 10 units); an unflagged
    value is ignored; the snapshot copy carries the result."""
    a, b, c = 2 * 32 + 2, 5 * 32 + 7, 9 * 32 + 9
    st = well_state(particles=[("GRVT", 10, 10, 50), ("GRVT", 38, 38, 50)])
    ov, se = st.arrays["grav_mass_override"], st.arrays["grav_mass_override_set"]
    ov[a], se[a] = 7, 1
    ov[b], se[b] = -123, 1
    ov[c], se[c] = 999, 0
    O.OracleState(st).call("oracle_step_gravity")
    m = st.arrays["grav_mass"]
    assert m[a] == 7 and m[b] == -123 and m[c] == 10 * 65536
    assert np.count_nonzero(m) == 3
    assert np.array_equal(st.arrays["grav_mass_prev"], m)
    assert ov[a] == 7 and se[a] == 1                       # the gravity step does not clear the channel


def test_tick_clears_the_override_channel_after_one_use():
    """The whole tick clears the channel where the element pass would, so an
    assignment lands in exactly one snapshot."""
    st = S.make_scene("612x384", scale_particles=0.02)
    cell = 30 * st.xcells + 40
    st.arrays["grav_mass_override"][cell] = 5 * 65536
    st.arrays["grav_mass_override_set"][cell] = 1
    o = O.OracleState(st)
    o.call("oracle_step", 0)
    assert st.arrays["grav_mass_prev"][cell] == 5 * 65536
    assert not st.arrays["grav_mass_override"].any() and not st.arrays["grav_mass_override_set"].any()
    o.call("oracle_step", 1)
    assert st.arrays["grav_mass_prev"][cell] == 0
