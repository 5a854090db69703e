"""Independent pins for the oracle functions that no reference test covers.

Every expected value here is computed in this file, in numpy float32 or by
hand, from the reference's Rust text; the oracle is only the thing checked.
A transcription slip in the oracle (a dropped term, a swapped axis, a wrong
sign, a wrong scan bound) changes at least one asserted bit.

* velocity integration and the sub-pixel step
  (move_kernel/velocity.rs:98-111, 133-181, 1056-1090);
* the gas bounce X-only / Y-only / both (velocity.rs:755-928);
* the powder diagonal and its perpendicular variant picked by ``rdir``
  (liquid.rs:298-350, 460-475);
* the liquid vertical scan with ``rt = 30`` / ``10`` and its asymmetric
  window, and the perpendicular scan that replaces it under Newtonian
  gravity (liquid.rs:388-617, 630-787, 801-963);
* vorticity confinement and fans (air_kernel/advect.rs:10-42, 296-330);
* Boussinesq and legacy convection (air_kernel/heat.rs:196-253);
* the particle -> air deposit and its highest-slot rule R1
  (move_kernel/mod.rs:297-346);
* consumption by VOID / PVOD / BHOL in the commit (commit.rs:76-128, 400-420).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S

from test_oracle_move import box, pcg

F = np.float32
NO_MOVE = 0xFFFFFFFF
XR = 128


def phys(name):
    """The element's move_lut record (move_kernel/mod.rs:1534-1576)."""
    lut = E.move_lut()
    t = E.pt(name)
    m = lut[t * 12:t * 12 + 12]
    f = m[4:].view(np.float32)
    return {"fall": int(m[1]), "adv": f[0], "drag": f[1], "aloss": f[2], "loss": f[3],
            "coll": f[4], "grav": f[5], "diff": f[6]}


def hashes(seed, slot):
    """h1 = pcg(seed ^ slot*2654435769), h2 = pcg(h1) (velocity.rs:1058-1059)."""
    h1 = pcg(seed ^ ((slot * 2654435769) & 0xFFFFFFFF))
    return h1, pcg(h1)


def integrate(name, vx0, vy0, ax, ay, seed, slot, newton=False):
    """integrate_velocity (velocity.rs:160-181) under vertical gravity, in
    float32 with the reference's left-to-right evaluation."""
    p = phys(name)
    h1, h2 = hashes(seed, slot)
    k = F(2.0) / F(65535.0)
    rdx = F(h1 & 0xFFFF) * k - F(1.0)
    rdy = F(h2 & 0xFFFF) * k - F(1.0)
    loss = F(1.0) if newton else p["loss"]
    vx = F(vx0) * loss + p["adv"] * F(ax) + F(0.0) + p["diff"] * rdx
    vy = F(vy0) * loss + p["adv"] * F(ay) + p["grav"] + p["diff"] * rdy
    return F(vx), F(vy), h1, h2


def step_axis(v, h):
    a = -v if v < 0 else v
    r = F(h & 0xFFFF) * (F(1.0) / F(65535.0))
    if a >= F(1.0) or r < a:
        return 1 if v > 0 else (-1 if v < 0 else 0)
    return 0


def round_i32(v):
    return int(np.trunc(F(v) + F(0.5)))


def put(st, k, name, x, y, vx=0.0, vy=0.0, flags=0):
    a = st.arrays
    a["part_ty"][k], a["part_x"][k], a["part_y"][k] = E.pt(name), x, y
    a["part_vx"][k], a["part_vy"][k], a["part_temp"][k] = vx, vy, 300.0
    a["part_flags"][k] = flags
    st.slot_bound = max(st.slot_bound, k + 1)


def scene(parts, **cfg):
    """128x128 walled box; parts = [(slot, name, x, y, vx, vy, flags)]."""
    st = box([], **cfg)
    st.slot_bound = 1
    for p in parts:
        put(st, *p)
    S.rebuild_pmap(st)
    return st


def intent(st, seed):
    O.OracleState(st).call("oracle_velocity_intent", seed)
    return st


def bits(v):
    return np.float32(v).view(np.uint32)


# --- velocity integration and the sub-pixel step ----------------------------

FREE = [
    ("WTRV", 0.0, 0.0, 0.0, 0.0), ("WTRV", 0.3, -0.2, 0.5, -0.25), ("WTRV", 3.0, 1.0, 0.0, 0.0),
    ("DUST", 0.0, 0.0, 0.0, 0.0), ("DUST", 0.4, 0.3, -0.3, 0.2), ("DUST", -2.5, 0.0, 0.0, 0.0),
    ("WATR", 0.2, 0.5, 0.1, 0.1), ("FIRE", 0.1, -0.3, 0.4, 0.0), ("GAS", -0.6, 0.2, 0.0, 0.3),
]


@pytest.mark.parametrize("name,vx0,vy0,ax,ay", FREE)
@pytest.mark.parametrize("seed,slot", [(99, 0), (12345, 5), (0xDEADBEEF, 17)])
def test_free_particle_velocity_and_target_bits(name, vx0, vy0, ax, ay, seed, slot):
    """A lone particle in open air: v = v0*loss + adv*air + g + diff*(h*2/65535-1)
    bit for bit, and the pixel it claims is round(p + v), or the stochastic
    unit step when that rounds back onto its own pixel."""
    x0, y0 = 64, 64
    st = scene([(slot, name, x0, y0, vx0, vy0, 0)])
    cell = (y0 // 4) * 32 + x0 // 4
    st.arrays["air_vx"][cell], st.arrays["air_vy"][cell] = ax, ay
    intent(st, seed)
    vx, vy, h1, h2 = integrate(name, vx0, vy0, ax, ay, seed, slot)
    assert max(abs(vx), abs(vy)) < 8.0
    fx, fy = round_i32(F(x0) + vx), round_i32(F(y0) + vy)
    if (fx, fy) == (x0, y0):
        sx, sy = step_axis(vx, h1), step_axis(vy, h2)
        fx, fy = x0 + sx, y0 + sy
    want = NO_MOVE if (fx, fy) == (x0, y0) else fy * XR + fx
    assert st.arrays["intent_target"][slot] == want
    assert bits(st.arrays["part_vx"][slot]) == bits(vx)
    assert bits(st.arrays["part_vy"][slot]) == bits(vy)
    if want != NO_MOVE:
        key = st.arrays["cell_winner"][want]
        assert key == (E.weight_ranks()[E.pt(name)] << 23) | (slot + 1)


def test_step_axis_rate_and_sign():
    """|v| < 1 steps when (h & 0xFFFF)/65535 < |v|; |v| >= 1 always steps."""
    for h in (0, 1, 0x7FFF, 0x8000, 0xFFFE, 0xFFFF, 0x12345678):
        r = (h & 0xFFFF) / 65535.0
        for v in (0.0, 0.25, 0.5, 0.999, 1.0, 3.0):
            for sgn in (1, -1):
                want = 0 if v == 0.0 else (sgn if (v >= 1.0 or r < v) else 0)
                assert step_axis(F(sgn * v), h) == want


# --- gas bounce -------------------------------------------------------------

def gas_bounce_case(seed, blocked):
    """WTRV at (64,64) flying at ~(3,3) into a METL block: the first PlanMove
    sample (one ISTP step) is the block; the bounce probes (2x - bx, by),
    then (bx, 2y - by)."""
    vx, vy, _, _ = integrate("WTRV", 10.0, 10.0, 0.0, 0.0, seed, 0)
    mv = max(abs(vx), abs(vy))
    assert mv >= F(2.0)
    fx = round_i32(F(64.0) + vx * F(2.0) / mv)
    fy = round_i32(F(64.0) + vy * F(2.0) / mv)
    assert fx in (65, 66) and fy in (65, 66)
    parts = [(0, "WTRV", 64, 64, 10.0, 10.0, 0)]
    k = 1
    metl = [(x, y) for x in (65, 66, 67) for y in (65, 66, 67)]
    rx, ry = 128 - fx, 128 - fy
    if blocked in ("y", "both"):
        metl.append((rx, fy))
    if blocked == "both":
        metl.append((fx, ry))
    for (x, y) in metl:
        parts.append((k, "METL", x, y, 0.0, 0.0, 0))
        k += 1
    st = intent(scene(parts), seed)
    return st, vx, vy, fx, fy, rx, ry


@pytest.mark.parametrize("seed", [3, 77, 1234, 99991])
def test_gas_bounce_x_only(seed):
    st, vx, vy, fx, fy, rx, ry = gas_bounce_case(seed, "x")
    c = phys("WTRV")["coll"]
    assert st.arrays["intent_target"][0] == fy * XR + rx
    assert bits(st.arrays["part_vx"][0]) == bits(vx * c)
    assert bits(st.arrays["part_vy"][0]) == bits(vy)


@pytest.mark.parametrize("seed", [3, 77, 1234, 99991])
def test_gas_bounce_y_only(seed):
    st, vx, vy, fx, fy, rx, ry = gas_bounce_case(seed, "y")
    c = phys("WTRV")["coll"]
    assert st.arrays["intent_target"][0] == ry * XR + fx
    assert bits(st.arrays["part_vx"][0]) == bits(vx)
    assert bits(st.arrays["part_vy"][0]) == bits(vy * c)


@pytest.mark.parametrize("seed", [3, 77, 1234, 99991])
def test_gas_bounce_both_axes(seed):
    st, vx, vy, fx, fy, rx, ry = gas_bounce_case(seed, "both")
    c = phys("WTRV")["coll"]
    assert st.arrays["intent_target"][0] == NO_MOVE
    assert bits(st.arrays["part_vx"][0]) == bits(vx * c)
    assert bits(st.arrays["part_vy"][0]) == bits(vy * c)


# --- powder diagonal --------------------------------------------------------

def seeds_for_rdir(slot, want, n=2):
    out, s = [], 1
    while len(out) < n:
        _, h2 = hashes(s, slot)
        if (1 if (h2 & 2) == 0 else -1) == want:
            out.append(s)
        s += 1
    return out


DIAG_SEEDS = [(s, r) for r in (1, -1) for s in seeds_for_rdir(0, r)]


@pytest.mark.parametrize("seed,rdir", DIAG_SEEDS)
@pytest.mark.parametrize("case", ["diagonal", "perpendicular", "none"])
def test_powder_on_a_ledge_takes_the_rdir_diagonal(seed, rdir, case):
    """DUST falling (vy = 0.8 + 0.1) onto a one-pixel METL ledge: the forward
    and clear-axis probes fail, so it takes (x - rdir, y + 1), else the
    perpendicular variant (x + rdir, y + 1), with rdir = +1 iff h2 & 2 == 0;
    every fallback multiplies v by Collision (0 for DUST)."""
    vx, vy, _, _ = integrate("DUST", 0.0, 1.0, 0.0, 0.0, seed, 0)
    assert vx == F(0.0) and abs(vy - F(0.9)) < 1e-6
    parts = [(0, "DUST", 64, 64, 0.0, 1.0, 0), (1, "METL", 64, 65, 0.0, 0.0, 0)]
    if case in ("perpendicular", "none"):
        parts.append((2, "METL", 64 - rdir, 65, 0.0, 0.0, 0))
    if case == "none":
        parts.append((3, "METL", 64 + rdir, 65, 0.0, 0.0, 0))
    st = intent(scene(parts), seed)
    want = {"diagonal": 65 * XR + 64 - rdir, "perpendicular": 65 * XR + 64 + rdir, "none": NO_MOVE}[case]
    assert st.arrays["intent_target"][0] == want
    c = phys("DUST")["coll"]
    assert st.arrays["part_vx"][0] == vx * c and st.arrays["part_vy"][0] == vy * c


# --- liquid scans -----------------------------------------------------------

def pool_case(seed, rdir, run, stagnant, newton):
    """WATR P at (64,64) inside a pool (rows 63..65) that extends `run` pixels
    past P on the rdir side and one pixel on the other, on a METL floor
    (row 66).  All eight neighbours are WATR, so nt = 0 and no diagonal is
    tried; the scan looks for the first non-WATR pixel beside P."""
    parts = [(0, "WATR", 64, 64, 0.0, 1.0, 1 if stagnant else 0)]
    k = 1
    lo, hi = sorted((64 - 2 * rdir, 64 + rdir * run))
    for y in (63, 64, 65):
        for x in range(lo, hi + 1):
            if (x, y) != (64, 64):
                parts.append((k, "WATR", x, y, 0.0, 0.0, 0))
                k += 1
    for x in range(30, 100):
        parts.append((k, "METL", x, 66, 0.0, 0.0, 0))
        k += 1
    st = scene(parts, newton_grav=1 if newton else 0)
    return intent(st, seed)


POOL_SEEDS = [(s, r) for r in (1, -1) for s in seeds_for_rdir(0, r, 1)]


@pytest.mark.parametrize("seed,rdir", POOL_SEEDS)
@pytest.mark.parametrize("stagnant", [True, False])
def test_liquid_vertical_scan_window(seed, rdir, stagnant):
    """Vertical gravity without Newtonian gravity: the lateral scan starts at
    clear_x + rdir, runs at most rt steps inside [clear_x - rt, clear_x + rt)
    and lands on the first free pixel of row fin_y = 65; rt = 10 for a
    stagnant liquid with no foreign neighbour, else 30 (liquid.rs:374-381,
    630-733)."""
    rt = 10 if stagnant else 30
    # the window is asymmetric: +rdir reaches clear_x + rt - 1, -rdir reaches clear_x - rt
    last_ok = rt - 2 if rdir == 1 else rt - 1
    for run in (last_ok, last_ok + 1):
        st = pool_case(seed, rdir, run, stagnant, newton=False)
        j = 64 + rdir * (run + 1)
        want = 65 * XR + j if run <= last_ok else NO_MOVE
        assert st.arrays["intent_target"][0] == want, (run, rt)
        assert st.arrays["part_vx"][0] == 0.0 and st.arrays["part_vy"][0] == 0.0   # Collision = 0


@pytest.mark.parametrize("seed,rdir", POOL_SEEDS)
def test_liquid_uses_perpendicular_scan_under_newtonian_gravity(seed, rdir):
    """With newton_grav the vertical scan is skipped (liquid.rs:516) and the
    perpendicular scan runs: j counts 0..rt-1 along row clear_y, the same
    reach on both sides, then falls one pixel to row 65."""
    rt = 10
    for run in (rt - 1, rt):
        st = pool_case(seed, rdir, run, stagnant=True, newton=True)
        j = 64 + rdir * (run + 1)
        want = 65 * XR + j if run <= rt - 1 else NO_MOVE
        assert st.arrays["intent_target"][0] == want, run
    # the +rdir side reaches one pixel further than the vertical scan would
    if rdir == 1:
        assert pool_case(seed, 1, rt - 1, True, newton=False).arrays["intent_target"][0] == NO_MOVE


# --- vorticity confinement and fans -----------------------------------------

def rotation_state(omega, coeff):
    c = S.default_config(128, 128, capacity=64)
    c["vorticity_coeff"] = coeff
    st = S.SimState(c)
    n = 32
    y, x = np.mgrid[0:n, 0:n]
    st.arrays["air_ovx"][:] = (F(-omega) * (y - 16).astype(np.float32)).ravel()
    st.arrays["air_ovy"][:] = (F(omega) * (x - 16).astype(np.float32)).ravel()
    return st


def smooth_advect(st):
    O.OracleState(st).call("oracle_air_smooth_advect")
    return st.arrays["air_vx"].reshape(32, 32).copy(), st.arrays["air_vy"].reshape(32, 32).copy()


def test_vorticity_confinement_closed_form():
    """Solid-body rotation v = omega x r has vorticity_bmap = 2*omega in the
    window x, y in [2, n-3] and 0 outside it.  Inside, |w| is flat and the
    force vanishes; on the window's edge one neighbour reads 0, so
    grad|w| = omega/2*2 along that axis and the force is
    k * (dwy, -dwx) / (|grad| + 0.001) * w with k = coeff/5."""
    omega, coeff = 2.0 ** -7, 0.5
    ax0, ay0 = smooth_advect(rotation_state(omega, 0.0))
    ax1, ay1 = smooth_advect(rotation_state(omega, coeff))
    dvx, dvy = (ax1 - ax0).astype(np.float64), (ay1 - ay0).astype(np.float64)
    k, w = coeff / 5.0, 2.0 * omega

    def force(dwx, dwy):
        norm = np.hypot(dwx, dwy)
        return k * dwy / (norm + 0.001) * w, k * -dwx / (norm + 0.001) * w

    n = 32
    for y in range(n):
        for x in range(n):
            if not (2 <= x <= n - 3 and 2 <= y <= n - 3):
                assert dvx[y, x] == 0.0 and dvy[y, x] == 0.0
                continue
            dwx = (0.0 if x == n - 3 else w) - (0.0 if x == 2 else w)
            dwy = (0.0 if y == n - 3 else w) - (0.0 if y == 2 else w)
            fx, fy = force(0.5 * dwx, 0.5 * dwy)
            assert dvx[y, x] == pytest.approx(fx, rel=2e-4, abs=1e-9), (x, y)
            assert dvy[y, x] == pytest.approx(fy, rel=2e-4, abs=1e-9), (x, y)
    assert dvy[16, 2] < 0 < dvy[16, n - 3] and dvx[2, 16] > 0 > dvx[n - 3, 16]


def test_vorticity_general_field_against_numpy():
    """A smooth vortex with non-uniform vorticity: the confinement term from
    a float64 numpy restatement of advect.rs:10-42 and :296-322."""
    c = S.default_config(256, 256, capacity=64)
    c["vorticity_coeff"] = 1.0
    n = 64
    rng = np.random.default_rng(5)
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    r2 = (x - 30.5) ** 2 + (y - 33.0) ** 2
    g = np.exp(-r2 / 60.0)
    ovx = (-(y - 33.0) * 0.05 * g + 0.01 * rng.standard_normal((n, n))).astype(np.float32)
    ovy = ((x - 30.5) * 0.05 * g + 0.01 * rng.standard_normal((n, n))).astype(np.float32)
    block = np.zeros((n, n), np.uint32)
    block[40:44, 10:14] = 1

    def run(coeff):
        cc = dict(c)
        cc["vorticity_coeff"] = coeff
        st = S.SimState(cc)
        st.arrays["air_ovx"][:] = ovx.ravel()
        st.arrays["air_ovy"][:] = ovy.ravel()
        st.arrays["block_air"][:] = block.ravel()
        O.OracleState(st).call("oracle_air_smooth_advect")
        return st.arrays["air_vx"].reshape(n, n).astype(np.float64), st.arrays["air_vy"].reshape(n, n).astype(np.float64)

    vx0, vy0 = run(0.0)
    vx1, vy1 = run(1.0)
    X, Y = ovx.astype(np.float64), ovy.astype(np.float64)

    def vort(xx, yy):
        if not (xx > 1 and xx + 2 < n and yy > 1 and yy + 2 < n):
            return 0.0
        b = block[yy, xx] != 0
        d1 = 0.0 if (b or block[yy, xx + 1] or block[yy, xx - 1]) else Y[yy, xx + 1] - Y[yy, xx - 1]
        d2 = 0.0 if (b or block[yy + 1, xx] or block[yy - 1, xx]) else X[yy + 1, xx] - X[yy - 1, xx]
        return (d1 - d2) * 0.5

    k = 1.0 / 5.0
    worst = 0.0
    for yy in range(2, n - 2):
        for xx in range(2, n - 2):
            w = vort(xx, yy)
            dwx = (abs(vort(xx + 1, yy)) - abs(vort(xx - 1, yy))) * 0.5
            dwy = (abs(vort(xx, yy + 1)) - abs(vort(xx, yy - 1))) * 0.5
            norm = np.hypot(dwx, dwy)
            fx, fy = k * dwy / (norm + 0.001) * w, k * -dwx / (norm + 0.001) * w
            ex = abs((vx1[yy, xx] - vx0[yy, xx]) - fx)
            ey = abs((vy1[yy, xx] - vy0[yy, xx]) - fy)
            tol = 1e-6 + 1e-3 * max(abs(fx), abs(fy))
            assert ex <= tol and ey <= tol, (xx, yy, fx, fy)
            worst = max(worst, abs(fx), abs(fy))
    assert worst > 1e-3          # the term is exercised


def test_fan_cell_adds_its_velocity():
    """A WL_FAN (5) cell adds fan_vx/fan_vy after smoothing, advection and
    vorticity (advect.rs:324-328); other cells ignore their fan words."""
    def run(fan):
        st = rotation_state(2.0 ** -7, 0.5)
        st.arrays["wall_bmap"][10 * 32 + 9] = 5
        if fan:
            st.arrays["fan_vx"][10 * 32 + 9], st.arrays["fan_vy"][10 * 32 + 9] = 0.25, -0.5
            st.arrays["fan_vx"][20 * 32 + 20], st.arrays["fan_vy"][20 * 32 + 20] = 3.0, 3.0   # no fan wall
        return smooth_advect(st)
    ax0, ay0 = run(False)
    ax1, ay1 = run(True)
    assert ax1[10, 9] == F(ax0[10, 9]) + F(0.25) and ay1[10, 9] == F(ay0[10, 9]) + F(-0.5)
    mask = np.ones((32, 32), bool)
    mask[10, 9] = False
    assert np.array_equal(ax0[mask], ax1[mask]) and np.array_equal(ay0[mask], ay1[mask])


# --- convection ---------------------------------------------------------------

def convect(mode, dT, block_above=False, cell=(16, 16)):
    c = S.default_config(128, 128, capacity=64)
    c["convection_mode"] = mode
    c["aheat_enable"] = 1
    st = S.SimState(c)
    amb = F(c["ambient_air_temp"])
    st.arrays["air_hv"][:] = amb
    cx, cy = cell
    st.arrays["air_hv"][cy * 32 + cx] = amb + F(dT)
    if block_above:
        st.arrays["block_airh"][cy * 32 + cx] = 8
    O.OracleState(st).call("oracle_air_heat_convect")
    return st.arrays["air_vx"].reshape(32, 32), st.arrays["air_vy"].reshape(32, 32), amb


@pytest.mark.parametrize("dT", [50.0, 500.0, -100.0, 0.0])
def test_boussinesq_single_hot_cell(dT):
    """weight = min((hv - ambient)/1e4, 0.01) from the old hv, pushed along
    -gravity = (0, -1) under vertical gravity, at that cell only."""
    vx, vy, amb = convect(2, dT)
    hot = amb + F(dT)
    w = (hot - amb) / F(10000.0)
    w = min(w, F(0.01))
    want = np.zeros((32, 32), np.float32)
    want[16, 16] = F(0.0) + w * F(-1.0)
    assert np.array_equal(vy, want)
    assert np.all(vx == 0.0)


@pytest.mark.parametrize("dT", [50.0, 20.0, -40.0])
def test_legacy_convection_pushes_the_cell_below_a_hot_cell(dT):
    """Legacy: weight = ((hv - hv_left)*gx + (hv - hv_above)*gy)/5000 with
    g = (0, -1), applied only when positive.  A hot cell makes the cell
    below it rise by dT/5000 and leaves itself alone; a cold cell rises
    itself (its own hv - hv_above < 0 times gy = -1 is positive)."""
    vx, vy, amb = convect(1, dT)
    want = np.zeros((32, 32), np.float32)
    if dT > 0:
        # cell (16,17): (amb - hot) * -1 / 5000 > 0
        w = ((amb - (amb + F(dT))) * F(-1.0)) / F(5000.0)
        want[17, 16] = F(0.0) + w * F(-1.0)
    else:
        # the cold cell itself: (cold - amb) * -1 / 5000 > 0
        w = (((amb + F(dT)) - amb) * F(-1.0)) / F(5000.0)
        want[16, 16] = F(0.0) + w * F(-1.0)
    assert np.array_equal(vy, want)
    assert np.all(vx == 0.0)


def test_legacy_convection_respects_a_heat_blocked_cell_above():
    """The push needs the cell above not to block ambient heat (heat.rs:215)."""
    c = S.default_config(128, 128, capacity=64)
    c["convection_mode"] = 1
    st = S.SimState(c)
    amb = F(c["ambient_air_temp"])
    st.arrays["air_hv"][:] = amb
    st.arrays["air_hv"][16 * 32 + 16] = amb + F(50.0)
    st.arrays["block_airh"][16 * 32 + 16] = 8       # the hot cell blocks heat
    O.OracleState(st).call("oracle_air_heat_convect")
    assert np.all(st.arrays["air_vy"] == 0.0)


# --- air deposit (R1) ---------------------------------------------------------

def test_air_deposit_single_particle_and_highest_slot_rule():
    """air_v = air_v*AirLoss + AirDrag*v (mod.rs:344-345).  With several
    eligible particles in one cell, each reads the pre-pass air velocity and
    the highest slot's write is kept (rule R1); a frozen (stasis) or rigid
    particle writes nothing."""
    st = scene([(0, "DUST", 40, 40, 2.0, 1.0, 0),                 # cell (10,10), alone
                (3, "DUST", 80, 40, 1.5, -1.0, 0), (7, "WATR", 81, 42, -3.0, 0.5, 0),  # cell (20,10)
                (5, "DUST", 60, 60, 9.0, 9.0, 0x10)])              # rigid: skipped
    a = st.arrays
    c1, c2, c3 = 10 * 32 + 10, 10 * 32 + 20, 15 * 32 + 15
    a["air_vx"][[c1, c2, c3]] = [1.0, 0.5, 0.25]
    a["air_vy"][[c1, c2, c3]] = [-0.5, 2.0, 0.75]
    pre = (a["air_vx"].copy(), a["air_vy"].copy())
    O.OracleState(st).call("oracle_air_deposit")
    d, w = phys("DUST"), phys("WATR")
    assert a["air_vx"][c1] == pre[0][c1] * d["aloss"] + d["drag"] * F(2.0)
    assert a["air_vy"][c1] == pre[1][c1] * d["aloss"] + d["drag"] * F(1.0)
    assert a["air_vx"][c2] == pre[0][c2] * w["aloss"] + w["drag"] * F(-3.0)
    assert a["air_vy"][c2] == pre[1][c2] * w["aloss"] + w["drag"] * F(0.5)
    assert a["air_vx"][c3] == pre[0][c3] and a["air_vy"][c3] == pre[1][c3]
    others = np.ones(32 * 32, bool)
    others[[c1, c2]] = False
    assert np.array_equal(a["air_vx"][others], pre[0][others])


def test_air_deposit_skips_stasis_and_margin():
    st = scene([(0, "DUST", 40, 40, 2.0, 1.0, 0), (1, "DUST", 2, 2, 2.0, 1.0, 0)])
    st.arrays["wall_bmap"][10 * 32 + 10] = 18        # WL_STASIS, emap 0 < 8: frozen
    st.arrays["air_vx"][:] = 0.5
    O.OracleState(st).call("oracle_air_deposit")
    assert np.all(st.arrays["air_vx"] == F(0.5))


# --- commit: consumption by VOID / PVOD / BHOL --------------------------------

def commit_case(dest, dest_ctype=0, dest_tmp=0, dest_life=0, mover="DUST", dest_temp=300.0, mover_temp=400.0):
    st = scene([(0, mover, 40, 40, 0.0, 1.0, 1), (1, dest, 40, 41, 0.0, 0.0, 0)])
    a = st.arrays
    a["part_ctype"][1], a["part_tmp"][1], a["part_life"][1] = dest_ctype, dest_tmp, dest_life
    a["part_temp"][0], a["part_temp"][1] = mover_temp, dest_temp
    a["intent_target"][0] = 41 * XR + 40
    a["cell_winner"][41 * XR + 40] = 1
    O.OracleState(st).call("oracle_commit")
    return a


@pytest.mark.parametrize("ctype,tmp,eaten", [(0, 0, True), (E.pt("DUST"), 0, True), (E.pt("WATR"), 0, False),
                                             (E.pt("WATR"), 1, True), (E.pt("DUST"), 1, False)])
def test_void_consumes_what_its_filter_admits(ctype, tmp, eaten):
    """VOID admits everything with ctype 0, else exactly ctype, inverted by
    tmp & 1 (void_admits, commit.rs:145-155); an admitted mover is erased,
    a rejected one stays put (can = 0)."""
    assert E.build_can_move()[E.pt("DUST"), E.pt("VOID")] == 3
    a = commit_case("VOID", ctype, tmp)
    assert a["part_ty"][1] == E.pt("VOID") and a["part_y"][1] == 41.0
    if eaten:
        assert a["part_ty"][0] == 0
    else:
        assert a["part_ty"][0] == E.pt("DUST") and a["part_y"][0] == 40.0
    assert a["part_flags"][0] & 1 == 0           # a found destination clears FLAG_STAGNANT


@pytest.mark.parametrize("life,eaten", [(10, True), (9, False), (0, False)])
def test_pvod_consumes_only_when_powered(life, eaten):
    a = commit_case("PVOD", dest_life=life)
    assert (a["part_ty"][0] == 0) == eaten


def test_black_hole_swallows_and_takes_half_the_heat():
    """BHOL: the mover is erased and the hole gains half its temperature,
    clamped to [MIN_TEMP, MAX_TEMP] (commit.rs:404-411)."""
    a = commit_case("BHOL", dest_temp=300.0, mover_temp=401.0)
    assert a["part_ty"][0] == 0
    assert a["part_temp"][1] == F(300.0) + F(401.0) / F(2.0)
    a = commit_case("BHOL", dest_temp=9900.0, mover_temp=9000.0)
    assert a["part_temp"][1] == F(9999.0)


def test_commit_loser_and_non_winner_stay():
    """Only the pixel's winner moves (commit.rs:334-340)."""
    st = scene([(0, "DUST", 40, 40, 0.0, 1.0, 0), (1, "DUST", 60, 40, 0.0, 1.0, 0)])
    a = st.arrays
    a["intent_target"][0] = a["intent_target"][1] = 41 * XR + 50
    a["cell_winner"][41 * XR + 50] = (E.weight_ranks()[E.pt("DUST")] << 23) | 2
    O.OracleState(st).call("oracle_commit")
    assert (a["part_x"][0], a["part_y"][0]) == (40.0, 40.0)
    assert (a["part_x"][1], a["part_y"][1]) == (50.0, 41.0)
