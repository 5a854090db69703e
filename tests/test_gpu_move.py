"""HIP movement pass and whole tick vs the oracle (bit-exact integer state
and fields), including the 2048x2048 workload at full size through
size-independent invariants."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu
INT_FIELDS = ["part_ty", "part_x", "part_y", "part_life", "part_ctype", "part_tmp", "part_flags", "pmap",
              "photons", "intent_target", "cell_winner"]
FLOAT_FIELDS = ["part_vx", "part_vy", "part_temp", "air_vx", "air_vy", "air_pv", "air_hv"]


def compare(ref, got, names, label):
    """Integer arrays bit for bit; float arrays by value (+0 == -0: the sign of
    a zero produced by a zero-mass FFT carries no information)."""
    for n in names:
        a, b = ref.arrays[n], got.arrays[n]
        if a.dtype != np.float32:
            a, b = a.view(np.uint32), b.view(np.uint32)
        if not np.array_equal(a, b):
            bad = np.flatnonzero(a != b)
            raise AssertionError(f"{label}: {n} differs at {len(bad)} entries, first {bad[0]}: "
                                 f"oracle {ref.arrays[n][bad[0]]!r} hip {got.arrays[n][bad[0]]!r}")


@pytest.mark.parametrize("workload,scale,ticks,newton", [("612x384", 1.0, 10, True), ("612x384", 1.0, 10, False),
                                                          ("1600x900", 0.25, 4, True)])
def test_movement_bit_exact(gpu, workload, scale, ticks, newton):
    st = S.make_scene(workload, scale_particles=scale, newton_grav=newton)
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[move parity] {workload} newton={newton}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            seed = L.frame_seed(frame)
            ctx.recalc_free_particles(seed, False)
            o.call("oracle_recalc_free_particles", seed, 0)
            ctx.step_movement(seed)
            o.call("oracle_step_movement", seed)
        ctx.fetch(st)
    compare(ref, st, INT_FIELDS + FLOAT_FIELDS, f"{workload} after {ticks} ticks")


@pytest.mark.parametrize("workload,scale,ticks", [("612x384", 1.0, 20), ("1600x900", 0.5, 6), ("2048x2048", 0.1, 3)])
def test_whole_tick_bit_exact(gpu, workload, scale, ticks):
    st = S.make_scene(workload, scale_particles=scale)
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[tick parity] {workload} ticks={ticks}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(ticks):
            ctx.step(frame)
            o.call("oracle_step", frame)
        ctx.fetch(st)
    compare(ref, st, INT_FIELDS + FLOAT_FIELDS + ["block_air", "block_airh", "wall_emap", "grav_mass"],
            f"{workload} tick {ticks}")


def test_whole_tick_rare_branches(gpu):
    """Whole ticks over stacked pixels, INVIS/FILT, photons, dying FIRE, the
    element types with special conduction and off-pixel positions: the map
    rebuild of each movement pass feeds the next tick's incremental
    RecalcFreeParticles (with CheckStacking on frames 0 and 10).  Newtonian
    gravity is off: the holes CheckStacking makes carry mass, and the FFT
    field is held to a tolerance (the gravity tests), not bit for bit."""
    from edge_scenes import adversarial, off_pixel_positions, thermal_specials
    st = S.make_scene("612x384", scale_particles=0.5, newton_grav=False)
    rng = np.random.default_rng(31)
    adversarial(st, rng)
    thermal_specials(st, rng)
    off_pixel_positions(st, rng)
    ref = st.copy()
    o = O.OracleState(ref)
    print("[tick parity] rare branches, 12 ticks", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(12):
            ctx.step(frame)
            o.call("oracle_step", frame)
        ctx.fetch(st)
    compare(ref, st, INT_FIELDS + FLOAT_FIELDS + ["pmap_count", "part_equil", "block_airh", "wall_emap"],
            "rare branches tick 12")


def test_full_size_2048_invariants(gpu):
    """The 2048x2048 workload at full size (2.1M particles): the oracle is too
    slow for many ticks, so check what any correct tick preserves."""
    st = S.make_scene("2048x2048")
    n0 = int(np.count_nonzero(st.arrays["part_ty"][: st.slot_bound]))
    ty0 = st.arrays["part_ty"][: st.slot_bound].copy()
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(20):
            ctx.step(frame)
        ctx.fetch(st, ["part_ty", "part_x", "part_y", "pmap", "air_pv", "air_hv", "part_temp"])
    ty = st.arrays["part_ty"][: st.slot_bound]
    live = np.flatnonzero(ty)
    xs, ys = st.arrays["part_x"][live], st.arrays["part_y"][live]
    flat = ys.astype(np.int64) * 2048 + xs.astype(np.int64)
    gas = (E.element_props_table()[ty[live].astype(np.int64) * 32 + 1] & 0x10) != 0
    assert len(np.unique(flat[~gas])) == int((~gas).sum())        # one matter particle per pixel
    pm = st.arrays["pmap"]
    assert np.count_nonzero(pm) == int((~gas).sum())               # the map holds every particle once
    assert np.array_equal(np.sort((pm[pm != 0] >> 9) - 1), np.sort(live[~gas]).astype(np.uint32))
    # walled box: only the fires, their fuel (OIL, DUST) and the water that
    # puts fires out (steam condenses into water) can vanish
    assert len(live) <= n0
    lost = np.flatnonzero((ty0 != 0) & (ty[: len(ty0)] == 0))
    assert set(np.unique(ty0[lost]).tolist()) <= {E.pt(n) for n in ("FIRE", "OIL", "DUST", "WATR", "WTRV")}
    assert np.isfinite(st.arrays["air_pv"]).all() and np.abs(st.arrays["air_pv"]).max() <= 256
    t = st.arrays["part_temp"][live]
    assert t.min() >= 0 and t.max() <= 9999


def test_photon_refraction_bit_exact(gpu):
    """Glass blocks in a photon swarm: refraction, total internal reflection,
    dispersion draws and the glass walk, HIP vs the oracle."""
    st = S.make_scene("612x384", scale_particles=0.1, newton_grav=False)
    rng = np.random.default_rng(41)
    a = st.arrays
    occupied = set(zip(a["part_x"][: st.slot_bound].astype(int), a["part_y"][: st.slot_bound].astype(int)))
    base, add = st.slot_bound, []
    for _ in range(6):  # glass blocks, some with holes in their faces
        x0, y0 = int(rng.integers(30, 500)), int(rng.integers(30, 300))
        w, h = int(rng.integers(10, 60)), int(rng.integers(10, 50))
        for y in range(y0, y0 + h):
            for x in range(x0, x0 + w):
                if (x, y) not in occupied and rng.random() > 0.02:
                    occupied.add((x, y))
                    add.append(("GLAS" if rng.random() > 0.3 else "BGLA", x, y, 0.0, 0.0, 0))
    for _ in range(3000):
        x, y = int(rng.integers(10, 600)), int(rng.integers(10, 370))
        th = rng.uniform(0, 2 * np.pi)
        mask = int(rng.choice([0x3FFFFFFF, (1 << 14) | (1 << 15), 0x3FF << 10, 0]))
        add.append(("PHOT", x, y, 3.0 * np.cos(th), 3.0 * np.sin(th), mask))
    for j, (name, x, y, vx, vy, ctype) in enumerate(add):
        s = base + j
        a["part_ty"][s], a["part_x"][s], a["part_y"][s] = E.pt(name), x, y
        a["part_vx"][s], a["part_vy"][s], a["part_ctype"][s], a["part_temp"][s] = vx, vy, ctype, 300.0
        a["part_life"][s] = 680 if name == "PHOT" else 0   # PHOT's default life
    st.slot_bound = base + len(add)
    ref = st.copy()
    o = O.OracleState(ref)
    print("[move parity] photons and glass", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(8):
            seed = L.frame_seed(frame)
            ctx.recalc_free_particles(seed, False)
            o.call("oracle_recalc_free_particles", seed, 0)
            ctx.step_movement(seed)
            o.call("oracle_step_movement", seed)
        ctx.fetch(st)
    compare(ref, st, INT_FIELDS + FLOAT_FIELDS, "photons and glass")


def test_wire_walls_whole_tick_bit_exact(gpu):
    """Detector walls under falling particles, runs and columns of every
    conductor wall type, some charged at the start: WL_DETECT charging, the
    set_emap closure and emap decay through whole ticks, HIP vs the oracle."""
    st = S.make_scene("612x384", scale_particles=0.3, newton_grav=False)
    rng = np.random.default_rng(53)
    xc, yc = st.xcells, st.ycells
    bmap, emap = st.arrays["wall_bmap"], st.arrays["wall_emap"]
    wires = [3, 3, 3, 1, 2, 6, 5, 12, 18]
    for _ in range(40):
        w = int(rng.choice(wires))
        x0, y0 = int(rng.integers(2, xc - 20)), int(rng.integers(2, yc - 12))
        if rng.random() < 0.6:
            cells = [(x, y0) for x in range(x0, x0 + int(rng.integers(2, 18)))]
        else:
            cells = [(x0, y) for y in range(y0, y0 + int(rng.integers(2, 10)))]
        for (x, y) in cells:
            bmap[y * xc + x] = w
            if rng.random() < 0.05:
                emap[y * xc + x] = 16
    S.derive_block_masks(st)
    ref = st.copy()
    o = O.OracleState(ref)
    print("[tick parity] wire walls, 10 ticks", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(10):
            ctx.step(frame)
            o.call("oracle_step", frame)
        ctx.fetch(st)
    compare(ref, st, INT_FIELDS + FLOAT_FIELDS + ["wall_emap", "block_air", "block_airh"], "wire walls tick 10")


def long_wire_scene():
    """2048x2048 (512x512 cells): a serpentine of WALLELEC that climbs 480
    rows twice (one row per sweep upward), so the closure reaches its
    ycells + 2 sweep bound before it converges; ladder rungs; combs; random
    runs.  Each structure is fed by a WL_DETECT cell with a DUST particle on
    it: only cells charged this tick (emap 16) feed the closure, since the
    air step has decayed older charges by then."""
    st = S.make_scene("2048x2048", scale_particles=0.004, newton_grav=False, seed=9)
    xc, yc = st.xcells, st.ycells
    bmap = st.arrays["wall_bmap"].reshape(yc, xc)
    feeds = []
    bmap[20:500, 20] = 1          # column A, fed at its foot: climbs
    bmap[20, 20:41] = 1           # bridge at the top
    bmap[20:501, 40] = 1          # column B: falls in one sweep
    bmap[500, 40:61] = 1          # bridge at the bottom
    bmap[20:501, 60] = 1          # column C: climbs again
    feeds.append((20, 500))
    for y in range(100, 400, 2):  # a ladder of singleton rungs between three-wide caps
        bmap[y, 100] = 1
        bmap[y + 1, 99:102] = 1
    feeds.append((100, 400))
    bmap[300, 150:400] = 6        # a comb: spine with teeth above and below
    for x in range(152, 400, 6):
        bmap[260:300, x] = 6
        bmap[301:340, x + 3] = 6
    feeds.append((149, 300))
    rng = np.random.default_rng(71)
    for _ in range(300):
        w = int(rng.choice([3, 3, 1, 2, 6, 12]))
        x0, y0 = int(rng.integers(420, xc - 40)), int(rng.integers(2, yc - 40))
        if rng.random() < 0.6:
            bmap[y0, x0:x0 + int(rng.integers(2, 36))] = w
        else:
            bmap[y0:y0 + int(rng.integers(2, 36)), x0] = w
    a = st.arrays
    occupied = set(zip((a["part_x"][:st.slot_bound]).astype(int), (a["part_y"][:st.slot_bound]).astype(int)))
    k = st.slot_bound
    for (cx, cy) in feeds:
        bmap[cy, cx] = 3
        px, py = 4 * cx + 1, 4 * cy + 1
        assert (px, py) not in occupied
        a["part_ty"][k], a["part_x"][k], a["part_y"][k], a["part_temp"][k] = E.pt("DUST"), px, py, 300.0
        k += 1
    st.slot_bound = k
    S.derive_block_masks(st)
    S.rebuild_pmap(st)
    return st


def test_wire_closure_on_the_largest_grid(gpu):
    """The row-parallel closure against the serial oracle on 512x512 cells,
    including a wire that needs more sweeps than the bound allows."""
    st = long_wire_scene()
    ref = st.copy()
    o = O.OracleState(ref)
    print("[tick parity] 2048x2048 long wires, 3 ticks", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(3):
            ctx.step(frame)
            o.call("oracle_step", frame)
            print(f"  tick {frame} done", flush=True)
        ctx.fetch(st)
    em = ref.arrays["wall_emap"].reshape(512, 512)
    assert 0 < np.count_nonzero(em[20:501, 60]) < 481      # column C is partly charged: the bound stopped it
    compare(ref, st, INT_FIELDS + ["wall_emap", "block_air"], "long wires")
