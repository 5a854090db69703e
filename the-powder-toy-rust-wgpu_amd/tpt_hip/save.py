"""The ``.tptrs`` save format: decode, encode, and restore into a ``SimState``.

A ``.tptrs`` file (crates/tpt-save/src/lib.rs:3-8, 300-372) is the 8-byte
magic ``TPTRS001`` followed by a zlib stream of ``bincode 1.3.3`` applied to
``SaveData`` (lib.rs:64-101).  ``bincode::serialize`` / ``deserialize`` use
bincode's legacy options: little-endian fixed-width integers, ``u64`` lengths
in front of every ``Vec`` and ``String``, one byte for ``bool`` (0 or 1) and
for an ``Option`` tag, a ``u32`` variant index for a unit enum, struct fields
in declaration order with no framing, ``#[serde(skip)]`` fields absent, and
trailing bytes allowed.  That is restated here from bincode's published
format; the reference holds no ``.tptrs`` files, so the byte layout is pinned
against hand-assembled payloads and the reference's own round-trip tests
(lib.rs:719-840).

``restore`` follows ``SaveData::restore_into`` (lib.rs:242-293) for a fresh
engine: the config from ``SimulationConfig::new`` plus the fields the save
carries, live particles imported into consecutive slots from 0, the grids
fitted to the cell count, and the spatial maps rebuilt the way
``rebuild_spatial_maps`` does (buffers/spatial.rs:279-321).
"""

import struct
import zlib

import numpy as np

from . import elements as E
from . import scene as S

MAGIC = b"TPTRS001"           # lib.rs:38
SAVE_VERSION = 2              # lib.rs:47
MAX_CANVAS_DIM = 16384        # lib.rs:53
MIN_CANVAS_DIM = 64           # tpt-core/src/particle.rs:14
CELL = 4
DEFAULT_TEMP_K = np.float32(295.15)
MAX_PARTICLE_CAPACITY = 0x7FFFFF
MIN_PARTICLE_CAPACITY = 65536

PT_PLSM, PT_THDR, PT_INVIS, PT_FILT, PT_EMBR, PT_FIGH = 28, 96, 124, 134, 153, 159
RENDER_DEFAULT = 0x1 | 0x2 | 0x4 | 0x40

# SimulationConfig as serialized (tpt-core/src/particle.rs:166-275): name and
# bincode kind; `mod_elements_enabled` is `#[serde(skip)]` and absent.
# Kinds: u32, i32, f32, bool, enum (u32 variant index), opt_f32.
CONFIG_FIELDS = [
    ("xres", "u32"), ("yres", "u32"), ("paused", "bool"), ("aheat_enable", "bool"),
    ("pretty_powder", "bool"), ("edge_mode", "enum3"), ("gravity_mode", "enum4"),
    ("air_mode", "enum5"), ("ambient_air_temp", "f32"), ("edge_pressure", "f32"),
    ("display_mode", "u32"), ("colour_mode", "u32"), ("render_mode", "u32"),
    ("fire_intensity", "f32"), ("heat_enable", "bool"), ("newton_grav", "bool"),
    ("water_equal_test", "bool"), ("vorticity_coeff", "f32"), ("convection_mode", "u32"),
    ("edge_velocity_x", "f32"), ("edge_velocity_y", "f32"), ("custom_gravity_x", "f32"),
    ("custom_gravity_y", "f32"), ("temperature_scale", "u32"), ("decorations", "bool"),
    ("mouse_x", "i32"), ("mouse_y", "i32"), ("gravity_field_enabled", "bool"),
    ("gravity_zones_enabled", "bool"), ("grid", "bool"), ("grid_size", "i32"),
    ("grid_checkerboard", "bool"), ("hud", "bool"), ("debug_hud", "bool"),
    ("show_brush", "bool"), ("zoom_enabled", "bool"), ("zoom_window_x", "i32"),
    ("zoom_window_y", "i32"), ("zoom_factor", "i32"), ("zoom_scope_x", "i32"),
    ("zoom_scope_y", "i32"), ("zoom_scope_size", "i32"), ("heat_display_limit_min", "opt_f32"),
    ("heat_display_limit_max", "opt_f32"), ("heat_display_auto_x", "i32"),
    ("heat_display_auto_y", "i32"), ("heat_display_auto_w", "i32"), ("heat_display_auto_h", "i32"),
]

# GpuParticle (tpt-core/src/particle.rs:26-62): sixteen 4-byte fields, 64 bytes.
PARTICLE_DTYPE = np.dtype([
    ("ty", "<u4"), ("life", "<i4"), ("ctype", "<u4"), ("x", "<f4"), ("y", "<f4"), ("vx", "<f4"),
    ("vy", "<f4"), ("temp", "<f4"), ("flags", "<u4"), ("tmp", "<i4"), ("tmp2", "<i4"), ("tmp3", "<i4"),
    ("tmp4", "<i4"), ("dcolour", "<u4"), ("_pad0", "<u4"), ("_pad1", "<u4"),
])
assert PARTICLE_DTYPE.itemsize == 64

# RigidPixel (particle.rs:384-395), 36 bytes.
RIGID_PIXEL_DTYPE = np.dtype([
    ("local_x", "<f4"), ("local_y", "<f4"), ("elem_ty", "<u4"), ("temp", "<f4"), ("life", "<i4"),
    ("ctype", "<u4"), ("dcolour", "<u4"), ("grid_x", "<i4"), ("grid_y", "<i4"),
])
# RigidBody2D (particle.rs:397-413) before its pixel list; `materialised` is skipped.
RIGID_HEAD = [("id", "u32"), ("active", "bool"), ("cm_x", "f32"), ("cm_y", "f32"), ("vx", "f32"),
              ("vy", "f32"), ("angle", "f32"), ("omega", "f32"), ("mass", "f32"), ("inv_mass", "f32"),
              ("inertia", "f32"), ("inv_inertia", "f32")]

GRIDS_F32 = ["air_vx", "air_vy", "air_pv", "air_hv", "fan_vx", "fan_vy"]
GRIDS_U32 = ["wall_bmap", "wall_emap"]


class SaveError(ValueError):
    """A payload that ``SaveData::from_bytes`` would reject."""


def align_dim(v):
    """particle.rs:381-383: at least 64, rounded up to a multiple of CELL."""
    v = max(int(v), MIN_CANVAS_DIM)
    return (v + CELL - 1) // CELL * CELL


def default_sim_config(xres, yres):
    """SimulationConfig::new (particle.rs:290-346), every serialized field."""
    return {
        "xres": align_dim(xres), "yres": align_dim(yres), "paused": False, "aheat_enable": False,
        "pretty_powder": False, "edge_mode": 0, "gravity_mode": 0, "air_mode": 0,
        "ambient_air_temp": float(DEFAULT_TEMP_K), "edge_pressure": 0.0, "display_mode": 0,
        "colour_mode": 0, "render_mode": RENDER_DEFAULT, "fire_intensity": 1.0, "heat_enable": True,
        "newton_grav": False, "water_equal_test": False, "vorticity_coeff": float(np.float32(0.1)),
        "convection_mode": 2, "edge_velocity_x": 0.0, "edge_velocity_y": 0.0, "custom_gravity_x": 0.0,
        "custom_gravity_y": 1.0, "temperature_scale": 0, "decorations": True, "mouse_x": -1,
        "mouse_y": -1, "gravity_field_enabled": False, "gravity_zones_enabled": False, "grid": False,
        "grid_size": 1, "grid_checkerboard": False, "hud": True, "debug_hud": False, "show_brush": True,
        "zoom_enabled": False, "zoom_window_x": 0, "zoom_window_y": 0, "zoom_factor": 8,
        "zoom_scope_x": 0, "zoom_scope_y": 0, "zoom_scope_size": 32,
        "heat_display_limit_min": 0.0, "heat_display_limit_max": 9999.0, "heat_display_auto_x": 0,
        "heat_display_auto_y": 0, "heat_display_auto_w": align_dim(xres),
        "heat_display_auto_h": align_dim(yres),
    }


class SaveData:
    """One snapshot, field for field as ``SaveData`` (lib.rs:64-101)."""

    def __init__(self, version=SAVE_VERSION, config=None, particles=None, grids=None,
                 rigid_bodies=None, signs=None):
        self.version = version
        self.config = dict(config) if config is not None else default_sim_config(612, 384)
        self.particles = particles if particles is not None else np.zeros(0, PARTICLE_DTYPE)
        self.grids = {n: np.zeros(0, np.float32) for n in GRIDS_F32}
        self.grids.update({n: np.zeros(0, np.uint32) for n in GRIDS_U32})
        if grids:
            self.grids.update(grids)
        self.rigid_bodies = list(rigid_bodies or [])
        self.signs = list(signs or [])  # (x, y, text, ju)

    def retain_air_pressure(self, include_pressure):
        """lib.rs:224-240: without pressure the four air grids are dropped."""
        if not include_pressure:
            for n in ("air_vx", "air_vy", "air_pv", "air_hv"):
                self.grids[n] = np.zeros(0, np.float32)
        return self


# ---- bincode (legacy options) ----------------------------------------------

class _Reader:
    def __init__(self, buf):
        self.buf = memoryview(buf)
        self.pos = 0

    def take(self, n):
        if n < 0 or self.pos + n > len(self.buf):
            raise SaveError("failed to deserialize save data: io error: unexpected end of file")
        out = self.buf[self.pos:self.pos + n]
        self.pos += n
        return out

    def u8(self):
        return self.take(1)[0]

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def f32(self):
        return struct.unpack("<f", self.take(4))[0]

    def u64(self):
        return struct.unpack("<Q", self.take(8))[0]

    def boolean(self):
        v = self.u8()
        if v > 1:
            raise SaveError(f"failed to deserialize save data: invalid value: {v}, expected a boolean")
        return bool(v)

    def value(self, kind):
        if kind == "u32":
            return self.u32()
        if kind == "i32":
            return self.i32()
        if kind == "f32":
            return self.f32()
        if kind == "bool":
            return self.boolean()
        if kind.startswith("enum"):
            v, n = self.u32(), int(kind[4:])
            if v >= n:
                raise SaveError(f"failed to deserialize save data: invalid enum variant index {v}")
            return v
        if kind == "opt_f32":
            tag = self.u8()
            if tag > 1:
                raise SaveError(f"failed to deserialize save data: invalid Option tag {tag}")
            return self.f32() if tag else None
        raise AssertionError(kind)

    def array(self, dtype):
        n = self.u64()
        dt = np.dtype(dtype)
        if n > (len(self.buf) - self.pos) // dt.itemsize:
            raise SaveError("failed to deserialize save data: io error: unexpected end of file")
        return np.frombuffer(self.take(n * dt.itemsize), dtype=dt).copy()

    def string(self):
        raw = bytes(self.take(self.u64()))
        try:
            return raw.decode("utf-8")
        except UnicodeDecodeError as err:
            raise SaveError(f"failed to deserialize save data: invalid utf-8: {err}") from None


class _Writer:
    def __init__(self):
        self.parts = []

    def value(self, kind, v):
        if kind == "u32" or kind.startswith("enum"):
            self.parts.append(struct.pack("<I", int(v)))
        elif kind == "i32":
            self.parts.append(struct.pack("<i", int(v)))
        elif kind == "f32":
            self.parts.append(struct.pack("<f", float(v)))
        elif kind == "bool":
            self.parts.append(b"\x01" if v else b"\x00")
        elif kind == "opt_f32":
            self.parts.append(b"\x00" if v is None else b"\x01" + struct.pack("<f", float(v)))
        else:
            raise AssertionError(kind)

    def array(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=np.dtype(dtype))
        self.parts.append(struct.pack("<Q", a.shape[0]))
        self.parts.append(a.tobytes())

    def string(self, s):
        raw = s.encode("utf-8")
        self.parts.append(struct.pack("<Q", len(raw)) + raw)

    def bytes(self):
        return b"".join(self.parts)


def _read_body(r, with_signs):
    d = SaveData(version=r.u32())
    d.config = {name: r.value(kind) for name, kind in CONFIG_FIELDS}
    d.particles = r.array(PARTICLE_DTYPE)
    for n in GRIDS_F32:
        d.grids[n] = r.array("<f4")
    for n in GRIDS_U32:
        d.grids[n] = r.array("<u4")
    d.rigid_bodies = []
    for _ in range(r.u64()):
        body = {name: r.value(kind) for name, kind in RIGID_HEAD}
        body["pixels"] = r.array(RIGID_PIXEL_DTYPE)
        d.rigid_bodies.append(body)
    d.signs = []
    if with_signs:
        for _ in range(r.u64()):
            x, y = r.i32(), r.i32()
            text = r.string()
            d.signs.append((x, y, text, r.u8()))
    return d


def payload(data, with_signs=True):
    """bincode of ``data`` (``SaveData``; ``SaveDataV1`` without signs)."""
    w = _Writer()
    w.value("u32", data.version)
    for name, kind in CONFIG_FIELDS:
        w.value(kind, data.config[name])
    w.array(data.particles, PARTICLE_DTYPE)
    for n in GRIDS_F32:
        w.array(data.grids[n], "<f4")
    for n in GRIDS_U32:
        w.array(data.grids[n], "<u4")
    w.parts.append(struct.pack("<Q", len(data.rigid_bodies)))
    for body in data.rigid_bodies:
        for name, kind in RIGID_HEAD:
            w.value(kind, body[name])
        w.array(body["pixels"], RIGID_PIXEL_DTYPE)
    if with_signs:
        w.parts.append(struct.pack("<Q", len(data.signs)))
        for x, y, text, ju in data.signs:
            w.value("i32", x)
            w.value("i32", y)
            w.string(text)
            w.parts.append(struct.pack("<B", ju))
    return w.bytes()


def to_bytes(data, level=6):
    """SaveData::to_bytes (lib.rs:296-313): magic + zlib(bincode)."""
    return MAGIC + zlib.compress(payload(data, with_signs=data.version >= 2), level)


def from_bytes(raw):
    """SaveData::from_bytes (lib.rs:316-371), with its checks in its order."""
    raw = bytes(raw)
    if len(raw) <= len(MAGIC):
        raise SaveError("save data is too short to contain a header")
    if raw[:len(MAGIC)] != MAGIC:
        raise SaveError("not a Powder Toy (Rust) save file: bad magic header")
    try:
        body = zlib.decompress(raw[len(MAGIC):])
    except zlib.error as err:
        raise SaveError(f"failed to decompress save data: {err}") from None
    version = _Reader(body).u32()
    if version == 0 or version > SAVE_VERSION:
        raise SaveError(f"unsupported save file version {version} (this build supports 1..={SAVE_VERSION})")
    data = _read_body(_Reader(body), with_signs=version >= 2)
    xres, yres = data.config["xres"], data.config["yres"]
    if xres == 0 or yres == 0 or xres > MAX_CANVAS_DIM or yres > MAX_CANVAS_DIM:
        raise SaveError(f"save file has an invalid canvas size {xres}x{yres}")
    return data


def load(path):
    with open(path, "rb") as f:
        return from_bytes(f.read())


def save(data, path):
    with open(path, "wb") as f:
        f.write(to_bytes(data))


# ---- restore / capture -------------------------------------------------------

def _round_i32(v):
    """Rust ``f32::round() as i32``: half away from zero, saturating, NaN -> 0."""
    v = np.asarray(v, dtype=np.float64)
    r = np.where(np.isnan(v), 0.0, np.sign(v) * np.floor(np.abs(v) + 0.5))
    return np.clip(r, -2.0 ** 31, 2.0 ** 31 - 1).astype(np.int64)


def rebuild_spatial_maps(st):
    """buffers/spatial.rs:279-321: slots in ascending order; energy types go
    to ``photons`` (the last writer keeps the pixel), other matter to
    ``pmap`` unless an INVIS/FILT would cover an occupied pixel; every
    non-energy particle but THDR/EMBR/FIGH/PLSM adds to ``pmap_count``.
    Then the wall masks (spatial.rs:336-351)."""
    a = st.arrays
    xres, yres = st.cfg["xres"], st.cfg["yres"]
    for n in ("pmap", "photons", "pmap_count"):
        a[n][:] = 0
    bound = st.slot_bound
    ty = a["part_ty"][:bound].astype(np.int64)
    ids = np.nonzero(ty)[0]
    x, y = _round_i32(a["part_x"][ids]), _round_i32(a["part_y"][ids])
    ok = (x >= 0) & (y >= 0) & (x < xres) & (y < yres)
    ids, x, y = ids[ok], x[ok], y[ok]
    t = ty[ids]
    props = E.element_props_table()
    energy = np.zeros(t.shape, bool)
    known = t < E.PT_NUM
    energy[known] = (props[t[known] * E.ELEM_STRIDE + 1] & E.TYPE_ENERGY) != 0
    idx = y * xres + x
    packed = S.pmap_pack(ids, t)

    def first(sel):  # per pixel, the lowest slot among ``sel``
        u, at = np.unique(idx[sel], return_index=True)
        return u, packed[sel][at]

    def last(sel):  # per pixel, the highest slot among ``sel``
        u, at = np.unique(idx[sel][::-1], return_index=True)
        return u, packed[sel][::-1][at]

    # Slots are visited in ascending order: the highest energy slot keeps a
    # photons pixel; in pmap the highest covering slot wins, and INVIS/FILT
    # only fill a pixel nothing covers (then the first of them stays).
    u, v = last(energy)
    a["photons"][u] = v
    m = ~energy
    covers = m & (t != PT_INVIS) & (t != PT_FILT)
    u, v = first(m & ~covers)
    a["pmap"][u] = v
    u, v = last(covers)
    a["pmap"][u] = v
    counted = m & (t != PT_THDR) & (t != PT_EMBR) & (t != PT_FIGH) & (t != PT_PLSM)
    np.add.at(a["pmap_count"], idx[counted], 1)
    S.derive_block_masks(st)


def restore(data, grav_mass_mode=0):
    """SaveData::restore_into (lib.rs:242-293) into a fresh ``SimState``."""
    c = data.config
    base = default_sim_config(c["xres"], c["yres"])
    for k in ("paused", "aheat_enable", "pretty_powder", "edge_mode", "gravity_mode", "air_mode",
              "ambient_air_temp", "edge_pressure", "display_mode", "colour_mode"):
        base[k] = c[k]
    xres, yres = base["xres"], base["yres"]
    live = data.particles[data.particles["ty"] != 0]
    cap = S.default_config(xres, yres)["capacity"]
    # ensure_particle_capacity (layout.rs:607-621): double until it fits.
    while cap < min(len(live), MAX_PARTICLE_CAPACITY):
        cap *= 2
    cap = min(cap, MAX_PARTICLE_CAPACITY)
    cfg = S.default_config(xres, yres, capacity=cap)
    for k in ("aheat_enable", "heat_enable", "newton_grav", "water_equal_test"):
        cfg[k] = int(bool(base[k]))
    for k in ("edge_mode", "gravity_mode", "air_mode", "convection_mode"):
        cfg[k] = int(base[k])
    for k in ("ambient_air_temp", "edge_pressure", "edge_velocity_x", "edge_velocity_y",
              "custom_gravity_x", "custom_gravity_y", "vorticity_coeff"):
        cfg[k] = float(np.float32(base[k]))
    cfg["grav_mass_mode"] = grav_mass_mode
    st = S.SimState(cfg)
    st.sim_config = base
    a = st.arrays
    # clear_all_particles (particles.rs:1112-1142) defaults for free slots.
    a["part_temp"][:] = DEFAULT_TEMP_K
    a["part_equil"][:] = -1.0
    n = min(len(live), cap)
    live = live[:n]
    for f, dst in (("ty", "part_ty"), ("life", "part_life"), ("ctype", "part_ctype"), ("x", "part_x"),
                   ("y", "part_y"), ("vx", "part_vx"), ("vy", "part_vy"), ("temp", "part_temp"),
                   ("flags", "part_flags"), ("tmp", "part_tmp"), ("tmp2", "part_tmp2"), ("tmp3", "part_tmp3"),
                   ("tmp4", "part_tmp4"), ("dcolour", "part_dcolour")):
        a[dst][:n] = live[f]
    st.slot_bound = max(1, min(n, cap))  # used_slot_bound after set_particle_count(n)
    cells = st.xcells * st.ycells
    for g in GRIDS_F32 + GRIDS_U32:
        src = data.grids[g]
        a[g][:] = src if len(src) == cells else 0  # fit_f32 / fit_u32
    a["air_ovx"][:] = 0.0
    a["air_ovy"][:] = 0.0
    a["air_opv"][:] = 0.0
    a["air_ohv"][:] = np.float32(cfg["ambient_air_temp"])
    rebuild_spatial_maps(st)
    # mark_all_dirty flags the walls, so the next gravity step rebuilds the
    # propagation mask (gravity_kernel.rs:874-881).
    S.gravity_mask(st)
    st.rigid_bodies = data.rigid_bodies
    st.signs = data.signs
    return st


def capture(st, sim_config=None, include_pressure=True):
    """SaveData::capture_from (lib.rs:187-221): live slots below the slot
    bound in slot order, the grids verbatim."""
    a = st.arrays
    bound = min(st.slot_bound, len(a["part_ty"]))
    ids = np.nonzero(a["part_ty"][:bound])[0]
    p = np.zeros(len(ids), PARTICLE_DTYPE)
    for f, src in (("ty", "part_ty"), ("life", "part_life"), ("ctype", "part_ctype"), ("x", "part_x"),
                   ("y", "part_y"), ("vx", "part_vx"), ("vy", "part_vy"), ("temp", "part_temp"),
                   ("flags", "part_flags"), ("tmp", "part_tmp"), ("tmp2", "part_tmp2"), ("tmp3", "part_tmp3"),
                   ("tmp4", "part_tmp4"), ("dcolour", "part_dcolour")):
        p[f] = a[src][ids]
    cfg = st.cfg
    sc = dict(sim_config) if sim_config else getattr(st, "sim_config", None) or default_sim_config(
        cfg["xres"], cfg["yres"])
    sc = dict(sc)
    for k in ("aheat_enable", "heat_enable", "newton_grav", "water_equal_test"):
        sc[k] = bool(cfg[k])
    for k in ("edge_mode", "gravity_mode", "air_mode", "convection_mode"):
        sc[k] = int(cfg[k])
    for k in ("ambient_air_temp", "edge_pressure", "edge_velocity_x", "edge_velocity_y",
              "custom_gravity_x", "custom_gravity_y", "vorticity_coeff"):
        sc[k] = float(cfg[k])
    sc["xres"], sc["yres"] = cfg["xres"], cfg["yres"]
    grids = {g: a[g].copy() for g in GRIDS_F32 + GRIDS_U32}
    d = SaveData(SAVE_VERSION, sc, p, grids, getattr(st, "rigid_bodies", []), getattr(st, "signs", []))
    return d.retain_air_pressure(include_pressure)
