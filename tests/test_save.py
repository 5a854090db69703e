"""The .tptrs loader (crates/tpt-save/src/lib.rs): the reference's own
round-trip and rejection tests (lib.rs:719-840) restated, the bincode byte
layout pinned by a payload assembled by hand from bincode 1.3.3's published
encoding, and ``restore`` checked against ``SaveData::restore_into`` /
``rebuild_spatial_maps`` on small hand-built cases.  The reference ships no
.tptrs file, so no golden save exists to load."""
import struct
import zlib

import numpy as np
import pytest

from tpt_hip import save as V
from tpt_hip import scene as S


def sample_particle(x, y, ty):
    """lib.rs:662-681."""
    p = np.zeros(1, V.PARTICLE_DTYPE)
    p["ty"], p["life"], p["ctype"], p["x"], p["y"] = ty, 42, 2, x, y
    p["vx"], p["vy"], p["temp"], p["flags"] = 0.5, -0.25, 295.15, 0
    p["tmp"], p["tmp2"], p["tmp3"], p["tmp4"], p["dcolour"] = 1, 2, 3, 4, 0x00112233
    return p


def sample_save():
    """lib.rs:683-717: a 128x96 canvas, two particles, filled grids, two signs."""
    cfg = V.default_sim_config(128, 96)
    cells = (128 // 4) * (96 // 4)
    grids = {"air_vx": np.full(cells, 0.25, np.float32), "air_vy": np.full(cells, -0.5, np.float32),
             "air_pv": np.full(cells, 1.5, np.float32), "air_hv": np.full(cells, 295.15, np.float32),
             "fan_vx": np.zeros(cells, np.float32), "fan_vy": np.zeros(cells, np.float32),
             "wall_bmap": np.zeros(cells, np.uint32), "wall_emap": np.zeros(cells, np.uint32)}
    parts = np.concatenate([sample_particle(10.0, 20.0, 1), sample_particle(11.0, 21.0, 2)])
    return V.SaveData(V.SAVE_VERSION, cfg, parts, grids, [], [(32, 48, "Hot zone", 0), (90, 12, "Max 45 chars", 2)])


def roundtrip(d):
    return V.from_bytes(V.to_bytes(d))


def test_magic_header_is_written():
    assert V.to_bytes(sample_save())[:8] == b"TPTRS001"


def test_roundtrip_preserves_state():
    o = sample_save()
    r = roundtrip(o)
    assert r.version == V.SAVE_VERSION
    assert (r.config["xres"], r.config["yres"]) == (128, 96)
    assert len(r.particles) == 2 and r.particles[1]["ty"] == 2
    assert len(r.grids["air_pv"]) == 32 * 24
    assert r.config == o.config
    assert r.particles.tobytes() == o.particles.tobytes()
    for g in V.GRIDS_F32 + V.GRIDS_U32:
        assert np.array_equal(r.grids[g], o.grids[g])


def test_roundtrip_preserves_signs():
    r = roundtrip(sample_save())
    assert r.signs == [(32, 48, "Hot zone", 0), (90, 12, "Max 45 chars", 2)]


def test_version_one_payload_loads_without_signs():
    o = sample_save()
    o.version = 1
    raw = b"TPTRS001" + zlib.compress(V.payload(o, with_signs=False))
    r = V.from_bytes(raw)
    assert r.version == 1 and r.signs == []
    assert r.particles.tobytes() == o.particles.tobytes()
    assert np.array_equal(r.grids["air_pv"], o.grids["air_pv"])


def test_payload_with_and_without_pressure():
    o = sample_save()
    kept = sample_save().retain_air_pressure(True)
    for g in ("air_vx", "air_vy", "air_pv", "air_hv"):
        assert np.array_equal(kept.grids[g], o.grids[g])
    dropped = sample_save().retain_air_pressure(False)
    for g in ("air_vx", "air_vy", "air_pv", "air_hv"):
        assert len(dropped.grids[g]) == 0
    assert len(dropped.grids["wall_bmap"]) == 32 * 24
    assert np.array_equal(dropped.grids["fan_vx"], o.grids["fan_vx"])
    r = roundtrip(dropped)
    for g in ("air_vx", "air_vy", "air_pv", "air_hv"):
        assert len(r.grids[g]) == 0
    assert r.signs == dropped.signs


@pytest.mark.parametrize("mutate,msg", [
    (lambda b: b"X" + b[1:], "bad magic"),
    (lambda b: b[:5], "too short"),
    (lambda b: b[:-4], "decompress"),
])
def test_rejects_corrupt(mutate, msg):
    with pytest.raises(V.SaveError, match=msg):
        V.from_bytes(mutate(V.to_bytes(sample_save())))


@pytest.mark.parametrize("version", [0, V.SAVE_VERSION + 1])
def test_rejects_version(version):
    d = sample_save()
    d.version = version
    with pytest.raises(V.SaveError, match="unsupported save file version"):
        V.from_bytes(V.to_bytes(d))


@pytest.mark.parametrize("xres,yres", [(0, 96), (128, 0), (16388, 96), (128, 20000)])
def test_rejects_canvas_size(xres, yres):
    d = sample_save()
    d.config["xres"], d.config["yres"] = xres, yres
    with pytest.raises(V.SaveError, match="invalid canvas size"):
        V.from_bytes(V.to_bytes(d))


def test_rejects_truncated_body_and_bad_bool():
    body = V.payload(sample_save())
    with pytest.raises(V.SaveError, match="unexpected end"):
        V.from_bytes(b"TPTRS001" + zlib.compress(body[:-3]))
    bad = bytearray(body)
    bad[4 + 8] = 2  # `paused`, the first bool after version, xres, yres
    with pytest.raises(V.SaveError, match="boolean"):
        V.from_bytes(b"TPTRS001" + zlib.compress(bytes(bad)))


def test_payload_is_zlib_compressed():
    raw = V.to_bytes(sample_save())
    assert raw[8] == 0x78
    assert len(raw) < len(V.payload(sample_save()))


def test_byte_layout_by_hand():
    """A payload written out field by field from the struct definitions
    (lib.rs:64-101, particle.rs:26-62, 166-275, 384-413, sign.rs:28-38)
    and bincode's legacy encoding, compared with the encoder's bytes and
    decoded back."""
    le = struct.pack
    b = le("<I", 2)                                   # version
    b += le("<II", 64, 72)                            # xres, yres
    b += bytes([1, 0, 1])                             # paused, aheat_enable, pretty_powder
    b += le("<III", 2, 3, 4)                          # edge Loop, gravity Custom, air NoUpdate
    b += le("<ff", 300.0, 1.5)                        # ambient_air_temp, edge_pressure
    b += le("<III", 5, 6, 7)                          # display, colour, render modes
    b += le("<f", 0.75)                               # fire_intensity
    b += bytes([0, 1, 1])                             # heat_enable, newton_grav, water_equal_test
    b += le("<fI", 0.2, 1)                            # vorticity_coeff, convection_mode
    b += le("<ffff", 0.1, -0.1, 0.5, -0.5)            # edge velocity, custom gravity
    b += le("<I", 1) + bytes([0])                     # temperature_scale, decorations
    b += le("<ii", -3, 9)                             # mouse
    b += bytes([1, 0, 1])                             # gravity_field, gravity_zones, grid
    b += le("<i", 4) + bytes([1, 0, 1, 0, 1])         # grid_size, checkerboard, hud, debug_hud, brush, zoom
    b += le("<iiiiii", 1, 2, 3, 4, 5, 6)              # zoom window x/y, factor, scope x/y/size
    b += bytes([1]) + le("<f", 10.0) + bytes([0])     # heat limits Some(10.0), None
    b += le("<iiii", 7, 8, 9, 10)                     # heat_display_auto x/y/w/h
    b += le("<Q", 1) + le("<IiIfffffIiiiiIII", 5, -7, 9, 1.5, 2.5, 0.25, -0.5, 400.0, 3, -1, -2, -3, -4,
                          0xAABBCCDD, 0, 0)       # one particle
    b += le("<Q", 2) + le("<ff", 1.0, 2.0)            # air_vx
    for _ in range(5):
        b += le("<Q", 0)                              # air_vy, air_pv, air_hv, fan_vx, fan_vy
    b += le("<Q", 1) + le("<I", 8)                    # wall_bmap
    b += le("<Q", 0)                                  # wall_emap
    b += le("<Q", 1) + le("<I", 11) + bytes([1]) + le("<ffffffffff", *range(10))  # rigid body head
    b += le("<Q", 1) + le("<ffIfiIIii", 0.5, -0.5, 9, 300.0, 3, 4, 5, 6, 7)      # its pixel
    b += le("<Q", 1) + le("<ii", 3, 4) + le("<Q", 2) + b"hi" + bytes([1])        # one sign

    d = V.from_bytes(b"TPTRS001" + zlib.compress(b))
    c = d.config
    assert (c["xres"], c["yres"], c["paused"], c["aheat_enable"], c["pretty_powder"]) == (64, 72, True, False, True)
    assert (c["edge_mode"], c["gravity_mode"], c["air_mode"]) == (2, 3, 4)
    assert (c["display_mode"], c["colour_mode"], c["render_mode"]) == (5, 6, 7)
    assert (c["heat_enable"], c["newton_grav"], c["water_equal_test"]) == (False, True, True)
    assert (c["mouse_x"], c["mouse_y"], c["grid_size"], c["zoom_scope_size"]) == (-3, 9, 4, 6)
    assert (c["heat_display_limit_min"], c["heat_display_limit_max"]) == (10.0, None)
    assert c["heat_display_auto_h"] == 10
    p = d.particles[0]
    assert (p["ty"], p["life"], p["ctype"], p["temp"], p["tmp4"], p["dcolour"]) == (5, -7, 9, 400.0, -4, 0xAABBCCDD)
    assert list(d.grids["air_vx"]) == [1.0, 2.0] and list(d.grids["wall_bmap"]) == [8]
    body = d.rigid_bodies[0]
    assert body["id"] == 11 and body["active"] and body["inv_inertia"] == 9.0
    assert body["pixels"][0]["elem_ty"] == 9 and body["pixels"][0]["grid_y"] == 7
    assert d.signs == [(3, 4, "hi", 1)]
    assert V.payload(d) == b


def _state_from(particles, xres=64, yres=64, walls=None):
    cfg = V.default_sim_config(xres, yres)
    cells = (xres // 4) * (yres // 4)
    grids = {g: np.zeros(cells, np.float32) for g in V.GRIDS_F32}
    grids.update({g: np.zeros(cells, np.uint32) for g in V.GRIDS_U32})
    if walls is not None:
        grids["wall_bmap"][:] = walls
    return V.restore(V.SaveData(2, cfg, np.concatenate(particles), grids))


def test_restore_slots_and_maps():
    """restore_into: live particles fill slots 0.. in file order (ty 0
    skipped); rebuild_spatial_maps rounds half away from zero, skips
    off-canvas particles, keeps an occupied pixel under INVIS/FILT, lets the
    higher slot take the pixel otherwise, and does not count PLSM."""
    ps = [sample_particle(2.5, 3.0, 1),      # slot 0 -> (3, 3)
          sample_particle(7.0, 7.0, 0),      # skipped, takes no slot
          sample_particle(3.4, 2.6, 124),    # slot 1 INVIS on (3, 3): keeps slot 0
          sample_particle(-0.6, 5.0, 1),     # slot 2 rounds to -1: off canvas
          sample_particle(9.5, 9.5, 28),     # slot 3 PLSM at (10, 10), not counted
          sample_particle(10.0, 10.0, 2),    # slot 4 takes (10, 10)
          sample_particle(20.0, 20.0, 134)]  # slot 5 FILT on an empty pixel
    st = _state_from(ps)
    a = st.arrays
    assert st.slot_bound == 6
    assert list(a["part_ty"][:7]) == [1, 124, 1, 28, 2, 134, 0]
    assert a["part_temp"][6] == np.float32(295.15) and a["part_equil"][0] == -1.0
    W = 64
    assert a["pmap"][3 * W + 3] == S.pmap_pack(0, 1)
    assert a["pmap_count"][3 * W + 3] == 2
    assert a["pmap"][10 * W + 10] == S.pmap_pack(4, 2)
    assert a["pmap_count"][10 * W + 10] == 1
    assert a["pmap"][20 * W + 20] == S.pmap_pack(5, 134)
    assert a["pmap"].astype(bool).sum() == 3 and a["pmap_count"].sum() == 4


def test_restore_config_and_grids():
    """Only the fields restore_into copies survive; the rest take
    SimulationConfig::new defaults.  A pressureless save loads calm air with
    air_ohv at ambient; walls derive the air/heat masks."""
    d = sample_save()
    d.config.update(aheat_enable=True, newton_grav=True, edge_mode=1, ambient_air_temp=300.0,
                    vorticity_coeff=0.5, edge_pressure=2.0)
    d.retain_air_pressure(False)
    d.grids["wall_bmap"][5] = 8
    d.grids["wall_bmap"][6] = 14
    st = V.restore(V.from_bytes(V.to_bytes(d)))
    cfg = st.cfg
    assert cfg["aheat_enable"] == 1 and cfg["edge_mode"] == 1 and cfg["ambient_air_temp"] == 300.0
    assert cfg["edge_pressure"] == 2.0
    assert cfg["newton_grav"] == 0 and cfg["vorticity_coeff"] == np.float32(0.1)
    a = st.arrays
    assert not a["air_pv"].any() and not a["air_vx"].any()
    assert (a["air_ohv"] == np.float32(300.0)).all()
    assert a["block_air"][5] == 1 and a["block_airh"][5] == 8
    assert a["block_air"][6] == 0 and a["block_airh"][6] == 8


def test_restore_normalises_dimensions():
    d = sample_save()
    d.config["xres"], d.config["yres"] = 30, 130
    st = V.restore(d)
    assert (st.cfg["xres"], st.cfg["yres"]) == (64, 132)
    assert not st.arrays["air_pv"].any()  # grids of the wrong size load as zeros


def test_scene_capture_restore_roundtrip():
    """A synthetic scene saved and loaded keeps every live particle's fields
    (in slot order) and every grid."""
    st = S.make_scene("612x384", scale_particles=0.25)
    d = V.capture(st)
    back = V.restore(V.from_bytes(V.to_bytes(d)))
    ids = np.nonzero(st.arrays["part_ty"][:st.slot_bound])[0]
    n = len(ids)
    assert back.slot_bound == n
    for f in ("part_ty", "part_life", "part_ctype", "part_x", "part_y", "part_vx", "part_vy", "part_temp",
              "part_flags", "part_tmp", "part_tmp2", "part_tmp3", "part_tmp4", "part_dcolour"):
        assert st.arrays[f][ids].tobytes() == back.arrays[f][:n].tobytes(), f
    for g in V.GRIDS_F32 + V.GRIDS_U32:
        assert np.array_equal(st.arrays[g], back.arrays[g]), g
    assert back.arrays["pmap"].astype(bool).sum() == st.arrays["pmap"].astype(bool).sum()


def test_gravity_mask_flood_fill():
    """Cells walled off by WL_GRAV keep mask 0; WL_WALL does not stop the fill;
    a GRAV cell on the border is not a seed."""
    walls = np.zeros((16, 16), np.uint32)
    walls[4, 4:9] = walls[8, 4:9] = 14
    walls[4:9, 4] = walls[4:9, 8] = 14
    walls[10:14, 10:14] = 8
    walls[0, 0] = 14
    st = _state_from([sample_particle(1.0, 1.0, 1)], walls=walls.reshape(-1))
    m = st.arrays["grav_mask"].reshape(16, 16)
    inside = np.zeros((16, 16), bool)
    inside[5:8, 5:8] = True
    assert (m[inside] == 0).all()
    assert (m[walls == 14] == 0).all()
    assert (m[~inside & (walls != 14)] == 0xFFFFFFFF).all()
