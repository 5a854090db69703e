"""Pins the oracle's air step (oracle/air.c) with closed forms and invariants.

The reference has no kernel-level air tests (its air_kernel/ holds no #[test]),
so the oracle is pinned by what the update rules fix mathematically:
a single-cell impulse through the linear pressure/velocity updates
(air_kernel/pressure.rs:233, 286-287), the normalised 3x3 Gaussian
(advect.rs:215-225), fixed points, clamps (advect.rs:337-339) and the wall-mask
truth table (settle.rs:74-91).  Expected values are computed in float64 from
the physical constants, independently of the oracle's float32 code.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import scene as S

W1, W2 = math.exp(-2.0), math.exp(-4.0)
WSUM = 1.0 + 4 * W1 + 4 * W2


def blank(xres=160, yres=128, **cfg):
    c = S.default_config(xres, yres, capacity=1024)
    c.update(cfg)
    st = S.SimState(c)
    st.arrays["air_hv"][:] = np.float32(c["ambient_air_temp"])
    return st


def run(st, fname="oracle_step_air"):
    O.OracleState(st).call(fname)
    return st


def grid(st, name):
    return st.arrays[name].reshape(st.ycells, st.xcells)


def test_zero_state_is_fixed_point():
    st = run(blank(aheat_enable=1))
    assert not grid(st, "air_pv").any() and not grid(st, "air_vx").any() and not grid(st, "air_vy").any()
    # Smoothing a constant in float32 is exact only up to rounding of the
    # weighted sum: within one ulp-scale of ambient.
    assert np.allclose(st.arrays["air_hv"], 295.15, rtol=1e-6, atol=0)


def test_pressure_impulse_first_tick_closed_form():
    st = blank(vorticity_coeff=0.0)
    x0, y0, P = 20, 15, 1.0
    grid(st, "air_pv")[y0, x0] = P
    o = O.OracleState(st)
    o.call("oracle_air_pressure")
    opv = grid(st, "air_opv")
    assert opv[y0, x0] == pytest.approx(0.9999 * P, rel=1e-6)
    assert np.count_nonzero(opv) == 1
    o.call("oracle_air_velocity")
    ovx, ovy = grid(st, "air_ovx"), grid(st, "air_ovy")
    # v' = 0.2 * (p[x-1] - p[x+1]): air flows away from the impulse.
    assert ovx[y0, x0 + 1] == pytest.approx(0.2 * 0.9999 * P, rel=1e-6)
    assert ovx[y0, x0 - 1] == pytest.approx(-0.2 * 0.9999 * P, rel=1e-6)
    assert ovy[y0 + 1, x0] == pytest.approx(0.2 * 0.9999 * P, rel=1e-6)
    assert ovy[y0 - 1, x0] == pytest.approx(-0.2 * 0.9999 * P, rel=1e-6)
    assert ovx[y0 + 1, x0] == 0 and ovy[y0, x0 + 1] == 0
    o.call("oracle_air_smooth_advect")
    pv = grid(st, "air_pv")
    # Pressure is only smoothed: centre 1/WSUM, edge e^-2/WSUM, corner e^-4/WSUM.
    assert pv[y0, x0] == pytest.approx(0.9999 * P / WSUM, rel=1e-6)
    assert pv[y0, x0 + 1] == pytest.approx(0.9999 * P * W1 / WSUM, rel=1e-6)
    assert pv[y0 + 1, x0 + 1] == pytest.approx(0.9999 * P * W2 / WSUM, rel=1e-6)
    assert pv.sum() == pytest.approx(0.9999 * P, rel=1e-5)


def test_velocity_impulse_divergence_sign():
    st = blank()
    x0, y0, V = 20, 15, 2.0
    grid(st, "air_vx")[y0, x0] = V
    O.OracleState(st).call("oracle_air_pressure")
    opv = grid(st, "air_opv")
    # dp = vx[x-1] - vx[x+1]: compression ahead of the flow, rarefaction behind.
    assert opv[y0, x0 + 1] == pytest.approx(0.15 * V, rel=1e-6)
    assert opv[y0, x0 - 1] == pytest.approx(-0.15 * V, rel=1e-6)
    assert opv[y0, x0] == 0.0


def test_border_band_is_damped_toward_edge_pressure():
    st = blank(edge_pressure=0.0)
    grid(st, "air_pv")[:] = 10.0
    O.OracleState(st).call("oracle_air_pressure")
    opv = grid(st, "air_opv")
    assert opv[0, 10] == pytest.approx(8.0, rel=1e-6)          # ring: 0.8 mix only
    assert opv[1, 10] == pytest.approx(8.0 * 0.9999, rel=1e-6)  # band and interior: both
    assert opv[5, 10] == pytest.approx(10.0 * 0.9999, rel=1e-6)


def test_clamps():
    st = blank()
    grid(st, "air_pv")[10:14, 10:14] = 1000.0
    grid(st, "air_vx")[20:24, 20:24] = -900.0
    run(st)
    assert st.arrays["air_pv"].max() <= 256.0 and st.arrays["air_vx"].min() >= -256.0


def test_uniform_heat_stays_uniform():
    st = blank(aheat_enable=1, convection_mode=0)
    run(st)
    hv = st.arrays["air_hv"]
    assert np.allclose(hv, 295.15, rtol=1e-6, atol=0)


def test_wall_mask_truth_table():
    st = blank()
    wall = st.arrays["wall_bmap"]
    emap = st.arrays["wall_emap"]
    cases = {0: (0, 0), 8: (1, 8), 1: (1, 8), 16: (1, 8), 14: (0, 8), 5: (0, 0), 2: (1, 8)}
    for i, w in enumerate(cases):
        wall[i] = w
    wall[20] = 2          # powered EWALL: emap 5 decays to 4, still open
    emap[20] = 5
    st.arrays["elem_block_air"][30] = 1
    st.arrays["elem_block_airh"][30] = 8
    run(st, "oracle_air_block_masks")
    for i, (w, (ba, bh)) in enumerate(cases.items()):
        assert (st.arrays["block_air"][i], st.arrays["block_airh"][i]) == (ba, bh), w
    assert st.arrays["wall_emap"][20] == 4 and st.arrays["block_air"][20] == 0
    assert st.arrays["block_air"][30] == 1 and st.arrays["block_airh"][30] == 8


def test_far_advect_walk_stops_before_wall():
    xc, yc = 40, 30
    block = np.zeros(xc * yc, dtype=np.uint32)
    block[15 * xc + 14] = 1                 # wall 6 cells to the left of (20,15)
    tx, ty = ctypes.c_float(), ctypes.c_float()
    # Velocity +10 along x: backtrace 7 cells to the left, walked cell by cell.
    O.lib().oracle_far_advect_target(20, 15, 10.0, 0.0, block.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                     xc, yc, 0, ctypes.byref(tx), ctypes.byref(ty))
    assert (tx.value, ty.value) == (15.0, 15.0)   # stops on the cell before the wall
    block[:] = 0
    O.lib().oracle_far_advect_target(20, 15, 10.0, 0.0, block.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                     xc, yc, 0, ctypes.byref(tx), ctypes.byref(ty))
    assert tx.value == pytest.approx(20 - 7.0, abs=1e-5) and ty.value == 15.0


def test_hot_air_is_linear_in_particle_count():
    st = blank()
    wtrv = S.E.pt("WTRV")
    hot = float(np.float32(S.E.ELEMENTS[wtrv]["hot_air"]))
    pm = st.arrays["pmap"].reshape(st.cfg["yres"], st.cfg["xres"])
    pm[40, 40] = S.pmap_pack(0, wtrv)
    pm[41, 42] = S.pmap_pack(1, wtrv)
    pm[42, 41] = S.pmap_pack(2, wtrv)
    run(st, "oracle_hot_air_pressure")
    assert grid(st, "air_pv")[10, 10] == pytest.approx(3 * 4 * hot, rel=1e-6)
    assert np.count_nonzero(st.arrays["air_pv"]) == 1
