"""Pins the conduction / phase-transition oracle (oracle/thermal.c).

Closed forms from thermal_kernel.rs: every config element has heat capacity
1 (:373-386), so the equilibrium of a gated particle is the arithmetic mean
of itself and its conducting pmap neighbours (:466-503); the gate is
gpu_hash(seed ^ i*2654435761 ^ ty) % 250 < HeatConduct (:485-491), computed
here independently; block_airh saturates at bit 3 (:597-599); the apply pass
takes the highest conducting slot's equilibrium (:616-708); WATR boils at
373 K minus twice the cell pressure into WTRV and raises the cell pressure
by 0.5 (:855-897, :1119-1127); WTRV condenses into DSTW (:1008-1014)."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S

SALT = 2654435761


def gpu_hash(x):
    x = (x ^ 0x9E3779B9) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    return x ^ (x >> 16)


def gate(seed, i, ty):
    return gpu_hash(seed ^ ((i * SALT) & 0xFFFFFFFF) ^ ty) % 250 < E.ELEMENTS[ty]["heat_conduct"]


def state(parts, xres=64, yres=64, **cfg):
    c = S.default_config(xres, yres, capacity=256)
    c.update(cfg)
    st = S.SimState(c)
    st.arrays["air_hv"][:] = np.float32(295.15)
    for i, (name, x, y, temp) in enumerate(parts):
        a = st.arrays
        a["part_ty"][i], a["part_x"][i], a["part_y"][i], a["part_temp"][i] = E.pt(name), x, y, temp
    st.slot_bound = len(parts)
    S.rebuild_pmap(st)
    return st


def test_gpu_hash_known_answers():
    for v in (0, 1, 0xCAFEBABE, 0x9E3779B9):
        assert O.lib().oracle_gpu_hash(v) == gpu_hash(v)


def test_equilibrium_is_mean_of_conducting_block():
    temps = [300.0 + 7 * k for k in range(9)]
    parts = [("METL", 20 + k % 3, 20 + k // 3, temps[k]) for k in range(9)]
    # pick a seed where the centre (slot 4) passes the conduction gate
    seed = next(s for s in range(1000) if gate(s, 4, E.pt("METL")))
    st = state(parts, aheat_enable=0)
    O.OracleState(st).call("oracle_thermal_conduction", seed)
    assert st.arrays["part_equil"][4] == pytest.approx(np.mean(temps), rel=1e-6)
    for i in range(9):
        want_gated = gate(seed, i, E.pt("METL"))
        assert (st.arrays["part_equil"][i] >= 0) == want_gated


def test_uniform_temperature_is_fixed_point():
    parts = [("WATR", 10 + k % 5, 10 + k // 5, 310.0) for k in range(25)]
    st = state(parts, aheat_enable=1)
    st.arrays["air_hv"][:] = np.float32(310.0)
    O.OracleState(st).call("oracle_step_thermal", 12345, 0)
    assert np.all(st.arrays["part_temp"][:25] == np.float32(310.0))
    assert np.all(st.arrays["air_hv"] == np.float32(310.0))


def test_block_airh_counter_saturates():
    # 12 insulating-gate failures (heat_conduct 0 -> always fail) in one cell
    parts = [("INSL", 40 + k % 4, 40 + k // 4, 300.0) for k in range(12)]
    st = state(parts)
    ci = 10 * 16 + 10
    st.arrays["block_airh"][ci] = 5
    st.arrays["block_airh"][0] = 8
    O.OracleState(st).call("oracle_thermal_conduction", 1)
    assert st.arrays["block_airh"][ci] == 8          # 5 + 12 saturates at 8
    assert st.arrays["block_airh"][0] == 8           # untouched


def test_apply_takes_highest_slot_equilibrium():
    st = state([("METL", 20, 20, 300.0), ("METL", 21, 20, 400.0), ("METL", 22, 20, 500.0)])
    st.arrays["part_equil"][:3] = [310.0, 320.0, -1.0]
    O.OracleState(st).call("oracle_thermal_apply")
    # slot 0's neighbourhood holds slot 1 (gated): highest slot wins -> 320.
    assert list(st.arrays["part_temp"][:3]) == [320.0, 320.0, 320.0]


def test_water_boils_and_steam_condenses():
    st = state([("WATR", 20, 20, 380.0), ("WTRV", 40, 40, 300.0), ("WATR", 30, 30, 374.0)])
    st.arrays["part_equil"][:3] = [380.0, 300.0, 374.0]
    pv = st.arrays["air_pv"].reshape(16, 16)
    pv[7, 7] = 1.0                                     # WATR at 374 K under pv=1: 374-2 < 373, no boil
    O.OracleState(st).call("oracle_phase_transitions", 9, 0)
    ty = st.arrays["part_ty"]
    assert ty[0] == E.pt("WTRV") and st.arrays["part_life"][0] == 0
    assert pv[5, 5] == pytest.approx(0.5)              # liquid -> gas raises the cell pressure
    assert ty[1] == E.pt("DSTW")
    assert ty[2] == E.pt("WATR") and pv[7, 7] == 1.0


def test_convection_swaps_isolated_pair_only():
    # hot water below cold water swaps; a three-high chain swaps its lowest-slot edge only
    st = state([("WATR", 10, 11, 350.0), ("WATR", 10, 10, 300.0),
                ("WATR", 30, 12, 360.0), ("WATR", 30, 11, 340.0), ("WATR", 30, 10, 320.0)])
    st.arrays["part_equil"][:5] = -1.0
    O.OracleState(st).call("oracle_phase_transitions", 3, 0)
    t = st.arrays["part_temp"]
    assert (t[0], t[1]) == (300.0, 350.0)
    assert (t[2], t[3], t[4]) == (340.0, 360.0, 320.0)
