"""ctypes binding of libtpt_hip.so, the C ABI of include/tpt_hip.h.

This is the host-side mirror of the reference's engine interface
(``SimulationEngine::{new, step}`` and the per-pass ``step_*_gpu`` functions,
crates/tpt-sim/src/lib.rs:140, 303 and the pass modules).  It only marshals
arguments: every pass runs in the HIP library, and a missing library raises
instead of falling back to anything else.
"""

import ctypes
import os

import numpy as np

from . import scene as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtpt_hip.so")

TPT_OK, TPT_EINVAL, TPT_ENOMEM, TPT_EDEVICE, TPT_ESTATE = 0, -1, -2, -3, -4
BUF = {name: i for i, name in enumerate(S.BUFFER_ORDER)}
PASS_SLOTS = 64


class TptConfig(ctypes.Structure):
    _fields_ = [
        ("xres", ctypes.c_uint32), ("yres", ctypes.c_uint32), ("capacity", ctypes.c_uint32),
        ("aheat_enable", ctypes.c_uint32), ("heat_enable", ctypes.c_uint32),
        ("newton_grav", ctypes.c_uint32), ("edge_mode", ctypes.c_uint32),
        ("gravity_mode", ctypes.c_uint32), ("air_mode", ctypes.c_uint32),
        ("convection_mode", ctypes.c_uint32), ("ambient_air_temp", ctypes.c_float),
        ("edge_pressure", ctypes.c_float), ("edge_velocity_x", ctypes.c_float),
        ("edge_velocity_y", ctypes.c_float), ("custom_gravity_x", ctypes.c_float),
        ("custom_gravity_y", ctypes.c_float), ("vorticity_coeff", ctypes.c_float),
        ("grav_mass_mode", ctypes.c_uint32),
        ("water_equal_test", ctypes.c_uint32),
    ]


# Exported symbols and their ctypes signatures (include/tpt_hip.h).
_vp = ctypes.c_void_p
SIGNATURES = {
    "tpt_last_error": ([], ctypes.c_char_p),
    "tpt_create": ([ctypes.POINTER(TptConfig), ctypes.POINTER(_vp)], ctypes.c_int),
    "tpt_destroy": ([_vp], ctypes.c_int),
    "tpt_set_element_table": ([_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t], ctypes.c_int),
    "tpt_upload": ([_vp, ctypes.c_int, _vp, ctypes.c_size_t, ctypes.c_size_t], ctypes.c_int),
    "tpt_download": ([_vp, ctypes.c_int, _vp, ctypes.c_size_t, ctypes.c_size_t], ctypes.c_int),
    "tpt_buffer_bytes": ([_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "tpt_set_slot_bound": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_get_slot_bound": ([_vp, ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "tpt_step_air": ([_vp], ctypes.c_int),
    "tpt_step_gravity": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_gravity_solve": ([_vp], ctypes.c_int),
    "tpt_gravity_solve_noalloc": ([_vp, _vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp], ctypes.c_int),
    "tpt_recalc_free_particles": ([_vp, ctypes.c_uint32, ctypes.c_int], ctypes.c_int),
    "tpt_step_thermal": ([_vp, ctypes.c_uint32, ctypes.c_int], ctypes.c_int),
    "tpt_step_water_equal": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_step_reactions": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_step_movement": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_step": ([_vp, ctypes.c_uint64], ctypes.c_int),
    "tpt_run": ([_vp, ctypes.c_uint64, ctypes.c_uint32], ctypes.c_int),
    "tpt_batch_create": ([ctypes.c_uint32, _vp, ctypes.POINTER(_vp)], ctypes.c_int),
    "tpt_batch_destroy": ([_vp], ctypes.c_int),
    "tpt_batch_context": ([_vp, ctypes.c_uint32, ctypes.POINTER(_vp)], ctypes.c_int),
    "tpt_batch_step": ([_vp, ctypes.c_uint64, ctypes.c_uint32], ctypes.c_int),
    "tpt_batch_step_launches": ([_vp, ctypes.c_uint64, ctypes.c_uint32], ctypes.c_int),
    "tpt_batch_synchronize": ([_vp], ctypes.c_int),
    "tpt_batch_step_serial": ([_vp, ctypes.c_uint64, ctypes.c_uint32], ctypes.c_int),
    "tpt_batch_set_lanes": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_batch_autotune": ([_vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "tpt_batch_tuning": ([_vp, _vp, _vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)], ctypes.c_int),
    "tpt_sur_create": ([ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(_vp)], ctypes.c_int),
    "tpt_sur_destroy": ([_vp], ctypes.c_int),
    "tpt_sur_param_count": ([_vp, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "tpt_sur_init_params": ([_vp, ctypes.c_uint64], ctypes.c_int),
    "tpt_sur_set_params": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "tpt_sur_get_params": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "tpt_sur_get_grads": ([_vp, _vp, ctypes.c_size_t], ctypes.c_int),
    "tpt_sur_set_scales": ([_vp, _vp, _vp, ctypes.c_float], ctypes.c_int),
    "tpt_sur_forward_raw": ([_vp, _vp, _vp], ctypes.c_int),
    "tpt_sur_backward_raw": ([_vp, _vp], ctypes.c_int),
    "tpt_sur_scene": ([_vp, _vp, _vp], ctypes.c_int),
    "tpt_sur_plain_ticks": ([_vp, _vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_sur_dataset_alloc": ([_vp, ctypes.c_uint32, ctypes.c_uint32], ctypes.c_int),
    "tpt_sur_dataset_capture": ([_vp, _vp, ctypes.c_uint32, ctypes.c_uint32], ctypes.c_int),
    "tpt_sur_train": ([_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.POINTER(ctypes.c_float)],
                      ctypes.c_int),
    "tpt_sur_eval_loss": ([_vp, _vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)], ctypes.c_int),
    "tpt_sur_set_clip": ([_vp, ctypes.c_float], ctypes.c_int),
    "tpt_sur_set_rollout": ([_vp, ctypes.c_uint32], ctypes.c_int),
    "tpt_sur_set_fused": ([_vp, ctypes.c_int], ctypes.c_int),
    "tpt_sur_step_air": ([_vp, _vp, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "tpt_sur_last_check": ([_vp, _vp], ctypes.c_int),
    "tpt_frame_seed": ([ctypes.c_uint64], ctypes.c_uint32),
    "tpt_synchronize": ([_vp], ctypes.c_int),
    "tpt_set_profiling": ([_vp, ctypes.c_int], ctypes.c_int),
    "tpt_reset_pass_times": ([_vp], ctypes.c_int),
    "tpt_get_pass_times": ([_vp, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_uint64), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
                           ctypes.c_int),
}

_lib = None


class TptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tpt error {code}: {msg}")
        self.code = code


def load():
    """Load libtpt_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (argt, rest) in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = argt
            f.restype = rest
        _lib = lib
    return _lib


def _check(rc):
    if rc != TPT_OK:
        raise TptError(rc, load().tpt_last_error().decode())


def frame_seed(frame_count):
    return load().tpt_frame_seed(frame_count)


def _config(cfg):
    c = TptConfig()
    for name, _ in TptConfig._fields_:
        setattr(c, name, cfg[name])
    return c


class Context:
    """One tpt_ctx: device buffers for one canvas plus its HIP stream."""

    def __init__(self, cfg, handle=None):
        self.cfg = dict(cfg)
        self._owned = handle is None
        self._h = _vp()
        if handle is None:
            _check(load().tpt_create(ctypes.byref(_config(cfg)), ctypes.byref(self._h)))
        else:
            self._h = handle

    def close(self):
        if self._h and self._owned:
            _check(load().tpt_destroy(self._h))
        self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # buffers --------------------------------------------------------------
    def set_tables(self, props, lut):
        props = np.ascontiguousarray(props, dtype=np.uint32)
        lut = np.ascontiguousarray(lut, dtype=np.uint32)
        _check(load().tpt_set_element_table(self._h, props.ctypes.data, props.size, lut.ctypes.data, lut.size))

    def upload(self, name, arr, offset_words=0):
        arr = np.ascontiguousarray(arr)
        assert arr.dtype.itemsize == 4
        _check(load().tpt_upload(self._h, BUF[name], arr.ctypes.data, offset_words * 4, arr.nbytes))

    def download(self, name, dtype, count=None, offset_words=0):
        if count is None:
            nbytes = ctypes.c_size_t()
            _check(load().tpt_buffer_bytes(self._h, BUF[name], ctypes.byref(nbytes)))
            count = nbytes.value // 4 - offset_words
        out = np.empty(count, dtype=dtype)
        _check(load().tpt_download(self._h, BUF[name], out.ctypes.data, offset_words * 4, out.nbytes))
        return out

    def set_slot_bound(self, bound):
        _check(load().tpt_set_slot_bound(self._h, bound))

    def load_state(self, st):
        """Upload every array of a SimState (sync_host_to_gpu of all categories)."""
        self.set_tables(st.arrays["elem_props"], st.arrays["move_lut"])
        self.upload("photon_mask", st.arrays["photon_mask"])
        self.upload("spawn_defaults", st.arrays["spawn_defaults"])
        for name in S.BUFFER_ORDER:
            if name not in S.TABLES:
                self.upload(name, st.arrays[name])
        self.set_slot_bound(st.slot_bound)

    def slot_bound(self):
        out = ctypes.c_uint32()
        _check(load().tpt_get_slot_bound(self._h, ctypes.byref(out)))
        return out.value

    def fetch(self, st, names=None):
        """Download arrays into a SimState (sync_gpu_to_host), with the slot bound."""
        for name in names or [n for n in S.BUFFER_ORDER if n not in S.TABLES]:
            arr = st.arrays[name]
            arr[:] = self.download(name, arr.dtype, arr.size)
        st.slot_bound = self.slot_bound()
        return st

    # passes ---------------------------------------------------------------
    def step_air(self):
        _check(load().tpt_step_air(self._h))

    def step_gravity(self, seed):
        _check(load().tpt_step_gravity(self._h, seed))

    def gravity_solve(self):
        _check(load().tpt_gravity_solve(self._h))

    def recalc_free_particles(self, seed, stacking):
        _check(load().tpt_recalc_free_particles(self._h, seed, 1 if stacking else 0))

    def step_thermal(self, seed, legacy=False):
        _check(load().tpt_step_thermal(self._h, seed, 1 if legacy else 0))

    def step_movement(self, seed):
        _check(load().tpt_step_movement(self._h, seed))

    def step_reactions(self, seed):
        _check(load().tpt_step_reactions(self._h, seed))

    def step_water_equal(self, seed):
        _check(load().tpt_step_water_equal(self._h, seed))

    def step(self, frame_count):
        _check(load().tpt_step(self._h, frame_count))

    def run(self, first_frame, ticks):
        """`ticks` ticks from `first_frame`: one eager tick, then graph replays."""
        _check(load().tpt_run(self._h, first_frame, ticks))

    def synchronize(self):
        _check(load().tpt_synchronize(self._h))

    # timing ---------------------------------------------------------------
    def set_profiling(self, on):
        _check(load().tpt_set_profiling(self._h, 1 if on else 0))

    def reset_pass_times(self):
        _check(load().tpt_reset_pass_times(self._h))

    def pass_times(self):
        names = (ctypes.c_char_p * PASS_SLOTS)()
        ms = (ctypes.c_double * PASS_SLOTS)()
        n_l = (ctypes.c_uint64 * PASS_SLOTS)()
        n = ctypes.c_int()
        _check(load().tpt_get_pass_times(self._h, names, ms, n_l, PASS_SLOTS, ctypes.byref(n)))
        return {names[i].decode(): (ms[i], n_l[i]) for i in range(n.value)}


class Batch:
    """tpt_batch: independent canvases advanced together (one graph replay
    per tick); `contexts[k]` loads and reads canvas k."""

    def __init__(self, cfgs):
        arr = (TptConfig * len(cfgs))(*[_config(c) for c in cfgs])
        self._h = _vp()
        _check(load().tpt_batch_create(len(cfgs), arr, ctypes.byref(self._h)))
        self.contexts = []
        for k, c in enumerate(cfgs):
            h = _vp()
            _check(load().tpt_batch_context(self._h, k, ctypes.byref(h)))
            self.contexts.append(Context(c, handle=h))

    def step(self, first_frame, ticks):
        _check(load().tpt_batch_step(self._h, first_frame, ticks))

    def step_launches(self, first_frame, ticks):
        _check(load().tpt_batch_step_launches(self._h, first_frame, ticks))

    def step_serial(self, first_frame, ticks):
        _check(load().tpt_batch_step_serial(self._h, first_frame, ticks))

    def set_lanes(self, lanes):
        _check(load().tpt_batch_set_lanes(self._h, lanes))

    def autotune(self, first_frame, trial_ticks):
        """Returns (chosen lanes, [(lanes, ms per batch tick), ...])."""
        best = ctypes.c_uint32()
        _check(load().tpt_batch_autotune(self._h, first_frame, trial_ticks, ctypes.byref(best)))
        lanes = (ctypes.c_uint32 * 64)()
        ms = (ctypes.c_float * 64)()
        n = ctypes.c_uint32()
        _check(load().tpt_batch_tuning(self._h, lanes, ms, 64, ctypes.byref(n)))
        return best.value, [(int(lanes[i]), float(ms[i])) for i in range(min(n.value, 64))]

    def synchronize(self):
        _check(load().tpt_batch_synchronize(self._h))

    def close(self):
        if self._h:
            for c in self.contexts:
                c.close()
            _check(load().tpt_batch_destroy(self._h))
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
