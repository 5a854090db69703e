"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, the smoke check and the CPU-baseline leg of bench.py to
check the HIP library against; the product path never imports this module.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f32p = ctypes.POINTER(ctypes.c_float)

# (field, ctype) in the order of struct oracle_state (tpt_oracle.h).
_CFG = [
    ("xres", ctypes.c_uint32), ("yres", ctypes.c_uint32), ("capacity", ctypes.c_uint32),
    ("slot_bound", ctypes.c_uint32), ("aheat_enable", ctypes.c_uint32),
    ("heat_enable", ctypes.c_uint32), ("newton_grav", ctypes.c_uint32),
    ("edge_mode", ctypes.c_uint32), ("gravity_mode", ctypes.c_uint32),
    ("air_mode", ctypes.c_uint32), ("convection_mode", ctypes.c_uint32),
    ("grav_mass_mode", ctypes.c_uint32), ("water_equal_test", ctypes.c_uint32),
    ("ambient_air_temp", ctypes.c_float), ("edge_pressure", ctypes.c_float),
    ("edge_velocity_x", ctypes.c_float), ("edge_velocity_y", ctypes.c_float),
    ("custom_gravity_x", ctypes.c_float), ("custom_gravity_y", ctypes.c_float),
    ("vorticity_coeff", ctypes.c_float),
]
# (field, dtype, extent) -- extent: "slot" | "pixel" | "cell" | "props" | "lut"
ARRAYS = [
    ("part_ty", np.uint32, "slot"), ("part_life", np.int32, "slot"),
    ("part_ctype", np.uint32, "slot"), ("part_x", np.float32, "slot"),
    ("part_y", np.float32, "slot"), ("part_vx", np.float32, "slot"),
    ("part_vy", np.float32, "slot"), ("part_temp", np.float32, "slot"),
    ("part_equil", np.float32, "slot"), ("part_flags", np.uint32, "slot"),
    ("part_tmp", np.int32, "slot"), ("part_tmp2", np.int32, "slot"),
    ("part_tmp3", np.int32, "slot"), ("part_tmp4", np.int32, "slot"),
    ("part_dcolour", np.uint32, "slot"), ("intent_target", np.uint32, "slot"),
    ("pmap", np.uint32, "pixel"), ("photons", np.uint32, "pixel"),
    ("pmap_count", np.uint32, "pixel"), ("cell_winner", np.uint32, "pixel"),
    ("air_vx", np.float32, "cell"), ("air_vy", np.float32, "cell"),
    ("air_pv", np.float32, "cell"), ("air_hv", np.float32, "cell"),
    ("air_ovx", np.float32, "cell"), ("air_ovy", np.float32, "cell"),
    ("air_opv", np.float32, "cell"), ("air_ohv", np.float32, "cell"),
    ("fan_vx", np.float32, "cell"), ("fan_vy", np.float32, "cell"),
    ("wall_bmap", np.uint32, "cell"), ("wall_emap", np.uint32, "cell"),
    ("block_air", np.uint32, "cell"), ("block_airh", np.uint32, "cell"),
    ("elem_block_air", np.uint32, "cell"), ("elem_block_airh", np.uint32, "cell"),
    ("grav_mask", np.uint32, "cell"), ("grav_mass", np.int32, "cell"),
    ("grav_mass_prev", np.int32, "cell"), ("grav_fx", np.float32, "cell"),
    ("grav_fy", np.float32, "cell"),
    ("elem_props", np.uint32, "props"), ("move_lut", np.uint32, "lut"),
    ("photon_mask", np.uint32, "table"), ("spawn_defaults", np.uint32, "table"),
    ("grav_tool_mass", np.float32, "cell"), ("grav_mass_override", np.int32, "cell"),
    ("grav_mass_override_set", np.uint32, "cell"),
]
_PTR = {np.uint32: _u32p, np.int32: _i32p, np.float32: _f32p}


class OracleStateC(ctypes.Structure):
    _fields_ = _CFG + [(name, _PTR[dt]) for name, dt, _ in ARRAYS]


_lib = None


def lib():
    """Load liboracle.so (building it first is the job of __graft_entry__.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB_PATH)
        for name in STATE_FUNCS:
            f = getattr(_lib, name)
            f.argtypes = [ctypes.POINTER(OracleStateC)]
            f.restype = None
        _lib.oracle_far_advect_target.argtypes = [
            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_float, _u32p,
            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, _f32p, _f32p]
        _lib.oracle_far_advect_target.restype = None
        _lib.oracle_gravity_solve_f64.argtypes = [ctypes.POINTER(OracleStateC), _f64p, _f64p]
        _lib.oracle_gravity_solve_f64.restype = None
        _lib.oracle_fft_twiddle_f32.argtypes = [ctypes.c_uint32, _f32p, _f32p]
        _lib.oracle_fft_transform_f32.argtypes = [_f32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32, _f32p, _f32p, ctypes.c_uint32,
                                                  ctypes.c_float]
        _lib.oracle_gravity_mask.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        _lib.oracle_lowbias_hash.argtypes = [ctypes.c_uint32]
        _lib.oracle_lowbias_hash.restype = ctypes.c_uint32
        _lib.oracle_recalc_free_particles.argtypes = [ctypes.POINTER(OracleStateC), ctypes.c_uint32, ctypes.c_int]
        _lib.oracle_recalc_free_particles.restype = None
        _lib.oracle_pcg_hash.argtypes = [ctypes.c_uint32]
        _lib.oracle_pcg_hash.restype = ctypes.c_uint32
        _lib.oracle_hash_u32.argtypes = [ctypes.c_uint32]
        _lib.oracle_hash_u32.restype = ctypes.c_uint32
        _lib.oracle_interact_wavelengths.argtypes = [ctypes.c_int32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32]
        _lib.oracle_interact_wavelengths.restype = ctypes.c_uint32
        for name, extra in (("oracle_air_deposit", []), ("oracle_detect_charge", []), ("oracle_emap_closure", []),
                            ("oracle_velocity_intent", [ctypes.c_uint32]),
                            ("oracle_commit", []), ("oracle_rebuild_maps", []),
                            ("oracle_step_movement", [ctypes.c_uint32]), ("oracle_step", [ctypes.c_uint64]),
                            ("oracle_water_equal", [ctypes.c_uint32]),
                            ("oracle_life_pass", [ctypes.c_uint32]), ("oracle_combustion_pass", [ctypes.c_uint32]),
                            ("oracle_fluid_pass", [ctypes.c_uint32]), ("oracle_step_reactions", [ctypes.c_uint32])):
            f = getattr(_lib, name)
            f.argtypes = [ctypes.POINTER(OracleStateC)] + extra
            f.restype = None
        _lib.oracle_gpu_hash.argtypes = [ctypes.c_uint32]
        _lib.oracle_gpu_hash.restype = ctypes.c_uint32
        for name, extra in (("oracle_thermal_conduction", [ctypes.c_uint32]), ("oracle_thermal_apply", []),
                            ("oracle_phase_transitions", [ctypes.c_uint32, ctypes.c_int]),
                            ("oracle_refresh_pmap_types", []),
                            ("oracle_step_thermal", [ctypes.c_uint32, ctypes.c_int])):
            f = getattr(_lib, name)
            f.argtypes = [ctypes.POINTER(OracleStateC)] + extra
            f.restype = None
        _lib.oracle_set_threads.argtypes = [ctypes.c_int]
        _lib.oracle_set_threads.restype = None
        _lib.oracle_max_threads.argtypes = []
        _lib.oracle_max_threads.restype = ctypes.c_int
        for name, argt in _OPTIONAL.items():
            f = getattr(_lib, name, None)
            if f is not None:
                f.argtypes = argt[0]
                f.restype = argt[1]
    return _lib


_OPTIONAL = {}
_f64p = ctypes.POINTER(ctypes.c_double)

# Oracle entry points that take only the state.
STATE_FUNCS = [
    "oracle_step_air", "oracle_air_pressure", "oracle_air_velocity", "oracle_air_smooth_advect",
    "oracle_air_heat_convect", "oracle_air_copy_back", "oracle_air_block_masks",
    "oracle_hot_air_pressure", "oracle_gravity_solve", "oracle_build_gravity_mass",
    "oracle_apply_gravity", "oracle_step_gravity", "oracle_clear_grav_override",
]


def register(name, argtypes, restype=None):
    """Declare an oracle entry point added by another translation unit."""
    _OPTIONAL[name] = (argtypes, restype)
    if _lib is not None and hasattr(_lib, name):
        f = getattr(_lib, name)
        f.argtypes = argtypes
        f.restype = restype


class OracleState:
    """Host-side state (numpy arrays) plus the C view the oracle works on.

    ``state`` is a :class:`tpt_hip.scene.SimState`; its arrays are used in
    place, so the oracle mutates them.
    """

    def __init__(self, state):
        self.state = state
        c = OracleStateC()
        cfg = state.cfg
        for name, ct in _CFG:
            if name == "slot_bound":
                setattr(c, name, state.slot_bound)
            elif name == "capacity":
                setattr(c, name, cfg["capacity"])
            else:
                setattr(c, name, cfg[name])
        for name, dt, _ in ARRAYS:
            arr = state.arrays[name]
            assert arr.dtype == dt and arr.flags.c_contiguous, name
            setattr(c, name, arr.ctypes.data_as(_PTR[dt]))
        self.c = c

    def call(self, fname, *args):
        self.c.slot_bound = self.state.slot_bound
        out = getattr(lib(), fname)(ctypes.byref(self.c), *args)
        self.state.slot_bound = self.c.slot_bound  # a movement pass may raise it
        return out


def gravity_solve_f64(state):
    """FP64 gravity field of state.grav_mass_prev -> (fx, fy) float64 arrays."""
    o = OracleState(state)
    n = state.xcells * state.ycells
    fx, fy = np.zeros(n), np.zeros(n)
    lib().oracle_gravity_solve_f64(ctypes.byref(o.c), fx.ctypes.data_as(_f64p), fy.ctypes.data_as(_f64p))
    return fx, fy


def fft_f32(data, rows, length, inner, stride, sign, tw_bits):
    """The reference's radix-2 transform (fft.rs) on an interleaved float32 array, in place."""
    n = 1 << tw_bits
    tre, tim = np.zeros(max(n // 2, 1), np.float32), np.zeros(max(n // 2, 1), np.float32)
    lib().oracle_fft_twiddle_f32(n, tre.ctypes.data_as(_f32p), tim.ctypes.data_as(_f32p))
    lib().oracle_fft_transform_f32(data.ctypes.data_as(_f32p), rows, length, inner, stride,
                                   tre.ctypes.data_as(_f32p), tim.ctypes.data_as(_f32p), tw_bits, sign)
    return data


def gravity_mask(wall, xcells, ycells):
    wall = np.ascontiguousarray(wall, dtype=np.uint32)
    mask = np.zeros(xcells * ycells, dtype=np.uint32)
    lib().oracle_gravity_mask(wall.ctypes.data_as(_u32p), xcells, ycells, mask.ctypes.data_as(_u32p))
    return mask
