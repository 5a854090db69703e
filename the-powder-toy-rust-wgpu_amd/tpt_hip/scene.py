"""Host-side simulation state and the synthetic scene generator.

``SimState`` holds every SimBuffers array (crates/tpt-sim/src/buffers/mod.rs:81-393)
as a numpy array in the reference layout, plus the SimulationConfig fields the
tick reads.  ``make_scene`` builds the synthetic workloads the benchmark and
the parity tests run:

* particles on distinct random pixels of the interior ``[4, W-4) x [4, H-4)``
  minus wall cells (the 4-px margin avoids the movement pass's edge kill,
  move_kernel/velocity.rs:247-249);
* slots assigned in creation order the way ``allocate_particle`` does,
  skipping the device spawn window ``[1024, 17408)``
  (buffers/particles.rs:156-163, buffers/mod.rs:502-508);
* element defaults from the element table (``allocate_particle``,
  particles.rs:184-203), then the scene's temperature draw;
* a ``WL_WALL`` ring on the border cells, optional wall rectangles;
* pressure as a sum of Gaussian blobs, zero air velocity, ambient heat
  ``ambient_air_temp``.
"""

import numpy as np

from . import elements as E

GPU_SPAWN_START = 1024
GPU_SPAWN_END = 1024 + 16384
WL_WALL = 8

SLOT_ARRAYS = [
    ("part_ty", np.uint32), ("part_life", np.int32), ("part_ctype", np.uint32),
    ("part_x", np.float32), ("part_y", np.float32), ("part_vx", np.float32),
    ("part_vy", np.float32), ("part_temp", np.float32), ("part_equil", np.float32),
    ("part_flags", np.uint32), ("part_tmp", np.int32), ("part_tmp2", np.int32),
    ("part_tmp3", np.int32), ("part_tmp4", np.int32), ("part_dcolour", np.uint32),
    ("intent_target", np.uint32),
]
PIXEL_ARRAYS = [("pmap", np.uint32), ("photons", np.uint32), ("pmap_count", np.uint32),
                ("cell_winner", np.uint32)]
CELL_ARRAYS = [
    ("air_vx", np.float32), ("air_vy", np.float32), ("air_pv", np.float32),
    ("air_hv", np.float32), ("air_ovx", np.float32), ("air_ovy", np.float32),
    ("air_opv", np.float32), ("air_ohv", np.float32), ("fan_vx", np.float32),
    ("fan_vy", np.float32), ("wall_bmap", np.uint32), ("wall_emap", np.uint32),
    ("block_air", np.uint32), ("block_airh", np.uint32), ("elem_block_air", np.uint32),
    ("elem_block_airh", np.uint32), ("grav_mask", np.uint32), ("grav_mass", np.int32),
    ("grav_mass_prev", np.int32), ("grav_fx", np.float32), ("grav_fy", np.float32),
    ("grav_tool_mass", np.float32), ("grav_mass_override", np.int32), ("grav_mass_override_set", np.uint32),
]
# Buffer ids of include/tpt_hip.h, in enum order.
BUFFER_ORDER = [n for n, _ in SLOT_ARRAYS] + [n for n, _ in PIXEL_ARRAYS] + \
    [n for n, _ in CELL_ARRAYS] + ["elem_props", "move_lut", "photon_mask", "spawn_defaults"]
TABLES = ["elem_props", "move_lut", "photon_mask", "spawn_defaults"]


def default_config(xres, yres, capacity=None):
    """SimulationConfig::new defaults (crates/tpt-core/src/particle.rs:225-262)."""
    pixels = xres * yres
    if capacity is None:
        # layout.rs:350-352: clamp(2 * pixels, 65536, 0x7FFFFF)
        capacity = min(max(2 * pixels, 65536), 0x7FFFFF)
    return {
        "xres": xres, "yres": yres, "capacity": capacity,
        "aheat_enable": 0, "heat_enable": 1, "newton_grav": 0,
        "edge_mode": 0, "gravity_mode": 0, "air_mode": 0, "convection_mode": 2,
        "ambient_air_temp": 295.15, "edge_pressure": 0.0, "edge_velocity_x": 0.0,
        "edge_velocity_y": 0.0, "custom_gravity_x": 0.0, "custom_gravity_y": 1.0,
        "vorticity_coeff": 0.1, "grav_mass_mode": 0, "water_equal_test": 0,
    }


class SimState:
    """All SimBuffers arrays for one canvas, host side."""

    def __init__(self, cfg):
        self.cfg = dict(cfg)
        xres, yres = cfg["xres"], cfg["yres"]
        assert xres % 4 == 0 and yres % 4 == 0
        self.xcells, self.ycells = xres // 4, yres // 4
        cap = cfg["capacity"]
        self.arrays = {}
        for n, dt in SLOT_ARRAYS:
            self.arrays[n] = np.zeros(cap, dtype=dt)
        self.arrays["intent_target"][:] = 0xFFFFFFFF
        for n, dt in PIXEL_ARRAYS:
            self.arrays[n] = np.zeros(xres * yres, dtype=dt)
        for n, dt in CELL_ARRAYS:
            self.arrays[n] = np.zeros(self.xcells * self.ycells, dtype=dt)
        self.arrays["elem_props"] = E.element_props_table()
        self.arrays["move_lut"] = E.move_lut()
        self.arrays["photon_mask"] = E.photon_mask()
        self.arrays["spawn_defaults"] = E.spawn_defaults()
        self.slot_bound = 1

    def copy(self):
        s = SimState.__new__(SimState)
        s.cfg = dict(self.cfg)
        s.xcells, s.ycells = self.xcells, self.ycells
        s.arrays = {k: v.copy() for k, v in self.arrays.items()}
        s.slot_bound = self.slot_bound
        return s

    def __getitem__(self, name):
        return self.arrays[name]


def slot_ids(n):
    """The first ``n`` slots ``allocate_particle`` hands out on a fresh pool."""
    ids = np.arange(n, dtype=np.int64)
    return np.where(ids < GPU_SPAWN_START, ids, ids + (GPU_SPAWN_END - GPU_SPAWN_START))


def pmap_pack(slot, ty):
    """constants.rs:814-816: ((part_id + 1) << 9) | (type & 0x1FF)."""
    return ((np.asarray(slot, dtype=np.uint64) + 1) << 9 | (np.asarray(ty, dtype=np.uint64) & 0x1FF)).astype(np.uint32)


def rebuild_pmap(st):
    """Host mirror of the map scatter for freshly placed, non-overlapping matter."""
    pm = st.arrays["pmap"]
    pm[:] = 0
    ty = st.arrays["part_ty"][:st.slot_bound]
    live = np.nonzero(ty)[0]
    x = st.arrays["part_x"][live].astype(np.int64)
    y = st.arrays["part_y"][live].astype(np.int64)
    energy = (E.element_props_table()[ty[live].astype(np.int64) * E.ELEM_STRIDE + 1] & E.TYPE_ENERGY) != 0
    m = ~energy
    pm[y[m] * st.cfg["xres"] + x[m]] = pmap_pack(live[m], ty[live][m])
    ph = st.arrays["photons"]
    ph[:] = 0
    ph[y[~m] * st.cfg["xres"] + x[~m]] = pmap_pack(live[~m], ty[live][~m])


def wall_ring(st):
    w = st.arrays["wall_bmap"].reshape(st.ycells, st.xcells)
    w[0, :] = WL_WALL
    w[-1, :] = WL_WALL
    w[:, 0] = WL_WALL
    w[:, -1] = WL_WALL


def wall_rects(st, rng, count, lo=8, hi=128):
    """``count`` axis-aligned WL_WALL rectangles, sides U{lo..hi} px, snapped to cells."""
    w = st.arrays["wall_bmap"].reshape(st.ycells, st.xcells)
    for _ in range(count):
        rw = int(rng.integers(lo, hi + 1)) // 4
        rh = int(rng.integers(lo, hi + 1)) // 4
        cx = int(rng.integers(1, max(2, st.xcells - rw - 1)))
        cy = int(rng.integers(1, max(2, st.ycells - rh - 1)))
        w[cy:cy + max(rh, 1), cx:cx + max(rw, 1)] = WL_WALL


def derive_block_masks(st):
    """The wall masks the air pass would derive from wall_bmap (settle.rs:51-93)."""
    wall = st.arrays["wall_bmap"] & 0xFF
    emap = st.arrays["wall_emap"] & 0xFF
    blocked = ((wall == 8) | (wall == 1) | (wall == 16) | ((wall == 2) & (emap == 0))).astype(np.uint32)
    st.arrays["block_air"][:] = blocked | st.arrays["elem_block_air"]
    st.arrays["block_airh"][:] = np.where((blocked != 0) | (wall == 14), 8, 0).astype(np.uint32) | \
        st.arrays["elem_block_airh"]


def gravity_mask(st):
    """rebuild_gravity_mask (gravity_kernel.rs:658-710): cells reachable from
    the canvas border through 4-neighbours without entering a WL_GRAV (14)
    cell get 0xFFFFFFFF, every other cell 0."""
    from scipy import ndimage
    open_ = (st.arrays["wall_bmap"].reshape(st.ycells, st.xcells) & 0xFF) != 14
    labels, _ = ndimage.label(open_, structure=[[0, 1, 0], [1, 1, 1], [0, 1, 0]])
    border = np.unique(np.concatenate([labels[0], labels[-1], labels[:, 0], labels[:, -1]]))
    reach = np.isin(labels, border[border != 0])
    st.arrays["grav_mask"][:] = np.where(reach.reshape(-1), 0xFFFFFFFF, 0).astype(np.uint32)


def gaussian_pressure(st, rng, blobs=8, sigma=6.0, amp=20.0):
    yy, xx = np.mgrid[0:st.ycells, 0:st.xcells].astype(np.float64)
    pv = np.zeros((st.ycells, st.xcells))
    for _ in range(blobs):
        cx, cy = rng.uniform(0, st.xcells), rng.uniform(0, st.ycells)
        a = rng.uniform(-amp, amp)
        pv += a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * sigma * sigma))
    st.arrays["air_pv"][:] = pv.reshape(-1).astype(np.float32)


# Element mixes of the workloads: (name, fraction).
MIX_BASE = [("DUST", 0.40), ("WATR", 0.30), ("WTRV", 0.15), ("METL", 0.15)]


def place_particles(st, rng, n, mix=MIX_BASE, temp_range=(273.0, 373.0), patches=()):
    """Place ``n`` particles on distinct free interior pixels.

    ``patches`` is a list of (element name, count) placed as 16x16 squares
    before the random fill (FIRE/OIL in the 1600x900 workload)."""
    xres, yres = st.cfg["xres"], st.cfg["yres"]
    wall_px = np.repeat(np.repeat(st.arrays["wall_bmap"].reshape(st.ycells, st.xcells) != 0, 4, 0), 4, 1)
    free = np.zeros((yres, xres), dtype=bool)
    free[4:yres - 4, 4:xres - 4] = True
    free &= ~wall_px
    chosen_px, chosen_ty = [], []
    for name, count in patches:
        placed = 0
        while placed < count:
            x0 = int(rng.integers(4, xres - 20))
            y0 = int(rng.integers(4, yres - 20))
            blk = free[y0:y0 + 16, x0:x0 + 16]
            ys, xs = np.nonzero(blk)
            take = min(len(ys), count - placed)
            if take == 0:
                continue
            sel = rng.permutation(len(ys))[:take]
            idx = (ys[sel] + y0) * xres + xs[sel] + x0
            free.reshape(-1)[idx] = False
            chosen_px.append(idx)
            chosen_ty.append(np.full(take, E.pt(name), dtype=np.uint32))
            placed += take
    remaining = n - sum(len(c) for c in chosen_px)
    cand = np.flatnonzero(free.reshape(-1))
    remaining = max(0, min(remaining, len(cand)))
    pick = rng.choice(cand, size=remaining, replace=False)
    names = [m[0] for m in mix]
    probs = np.array([m[1] for m in mix], dtype=np.float64)
    probs /= probs.sum()
    tys = np.array([E.pt(nm) for nm in names], dtype=np.uint32)[rng.choice(len(names), size=remaining, p=probs)]
    chosen_px.append(pick)
    chosen_ty.append(tys)
    px = np.concatenate(chosen_px).astype(np.int64)
    ty = np.concatenate(chosen_ty).astype(np.uint32)
    total = len(px)
    slots = slot_ids(total)
    cap = st.cfg["capacity"]
    assert slots[-1] < cap if total else True, "pool too small for the scene"
    a = st.arrays
    props = E.ELEMENTS
    a["part_ty"][slots] = ty
    a["part_x"][slots] = (px % xres).astype(np.float32)
    a["part_y"][slots] = (px // xres).astype(np.float32)
    a["part_vx"][slots] = 0.0
    a["part_vy"][slots] = 0.0
    for field, src, dt in (("part_life", "default_life", np.int32), ("part_ctype", "default_ctype", np.uint32),
                           ("part_tmp", "default_tmp", np.int32), ("part_tmp2", "default_tmp2", np.int32),
                           ("part_tmp3", "default_tmp3", np.int32), ("part_tmp4", "default_tmp4", np.int32)):
        table = np.array([e[src] for e in props], dtype=np.int64).astype(dt)
        a[field][slots] = table[ty]
    a["part_temp"][slots] = rng.uniform(temp_range[0], temp_range[1], size=total).astype(np.float32)
    fire = ty == E.pt("FIRE")
    if fire.any():
        a["part_life"][slots[fire]] = (120 + rng.integers(0, 50, size=int(fire.sum()))).astype(np.int32)
    st.slot_bound = int(slots[-1]) + 1 if total else 1
    rebuild_pmap(st)
    return total


WORKLOADS = {
    # name: (xres, yres, seed, particles, wall rectangles, patches)
    "612x384": (612, 384, 0, 20000, 0, ()),
    "1600x900": (1600, 900, 1, 403200, 0, (("FIRE", 28800), ("OIL", 43200))),
    # the 1600x900 mix: FIRE and OIL patches of 2% and 3% of the pixels
    "2048x2048": (2048, 2048, 1, 2097152, 64, (("FIRE", 83886), ("OIL", 125829))),
}


def make_scene(workload="612x384", *, scale_particles=1.0, newton_grav=True, aheat=True,
               grav_mass_mode=0, capacity=None, xres=None, yres=None, seed=None, particles=None):
    """Build one of the named workloads (or a custom canvas)."""
    w = WORKLOADS.get(workload, (xres, yres, seed or 0, particles or 0, 0, ()))
    X, Y, sd, n, rects, patches = w
    X, Y = xres or X, yres or Y
    sd = sd if seed is None else seed
    n = n if particles is None else particles
    n = int(n * scale_particles)
    patches = tuple((nm, int(c * scale_particles)) for nm, c in patches)
    rng = np.random.default_rng(sd)
    cfg = default_config(X, Y, capacity)
    cfg["aheat_enable"] = 1 if aheat else 0
    cfg["newton_grav"] = 1 if newton_grav else 0
    cfg["grav_mass_mode"] = grav_mass_mode
    if capacity is None:
        need = int(slot_ids(max(n, 1))[-1]) + 1
        cfg["capacity"] = min(max(cfg["capacity"], need), 0x7FFFFF)
    st = SimState(cfg)
    wall_ring(st)
    if rects:
        wall_rects(st, rng, rects)
    derive_block_masks(st)
    gaussian_pressure(st, rng)
    st.arrays["air_hv"][:] = np.float32(cfg["ambient_air_temp"])
    place_particles(st, rng, n, patches=patches)
    st.arrays["grav_mask"][:] = 0xFFFFFFFF
    return st
