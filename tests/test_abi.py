"""The C-ABI library loads and exports every entry point include/tpt_hip.h declares.

CPU-only: no compute call is made (there may be no GPU here)."""
import ctypes
import os
import re

import pytest

from tpt_hip import lib as L
from tpt_hip import scene as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "tpt_hip.h")).read()
    return sorted(set(re.findall(r"TPT_API\s+[\w\s\*]*?\b(tpt_\w+)\s*\(", src)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("tpt_create", "tpt_step", "tpt_step_air", "tpt_step_gravity", "tpt_step_thermal",
              "tpt_step_movement", "tpt_recalc_free_particles", "tpt_upload", "tpt_download"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(declared_symbols()) == set(L.SIGNATURES)


def test_frame_seed_closed_form():
    # lib.rs:315-318: rotl(frame*0x9E3779B9, 13) ^ 0x85EBCA6B, no device needed.
    def seed(fc):
        v = (fc * 0x9E3779B9) & 0xFFFFFFFF
        v = ((v << 13) | (v >> 19)) & 0xFFFFFFFF
        return v ^ 0x85EBCA6B
    for fc in (0, 1, 2, 9, 10, 12345, 2**32 + 7):
        assert L.frame_seed(fc) == seed(fc)
    assert L.frame_seed(0) == 0x85EBCA6B


def test_config_struct_matches_header():
    src = open(os.path.join(ROOT, "include", "tpt_hip.h")).read()
    body = re.search(r"typedef struct tpt_config \{(.*?)\} tpt_config;", src, re.S).group(1)
    fields = re.findall(r"(?:uint32_t|float)\s+(\w+);", body)
    assert fields == [f for f, _ in L.TptConfig._fields_]


def _create(**over):
    cfg = S.default_config(64, 64)
    cfg.update(over)
    return L.Context(cfg)


@pytest.mark.parametrize("over", [dict(xres=66), dict(yres=2), dict(xres=2900), dict(yres=0), dict(capacity=0),
                                  dict(capacity=0x800000), dict(edge_mode=3), dict(air_mode=5),
                                  dict(gravity_mode=4), dict(convection_mode=3), dict(grav_mass_mode=2)])
def test_create_rejects_bad_configs(over):
    """Validation happens before any device call, so this runs without a GPU:
    the reference silently no-ops on such canvases, the ABI says why."""
    with pytest.raises(L.TptError) as e:
        _create(**over)
    assert e.value.code == L.TPT_EINVAL
    assert "tpt error -1: " in str(e.value) and len(str(e.value)) > len("tpt error -1: ")


def test_null_arguments_are_rejected():
    lib = L.load()
    assert lib.tpt_step(None, 0) == L.TPT_EINVAL
    assert lib.tpt_step_movement(None, 1) == L.TPT_EINVAL
    assert lib.tpt_upload(None, 0, None, 0, 0) == L.TPT_EINVAL
    assert lib.tpt_download(None, 0, None, 0, 0) == L.TPT_EINVAL
    assert lib.tpt_set_slot_bound(None, 1) == L.TPT_EINVAL
    assert lib.tpt_get_slot_bound(None, None) == L.TPT_EINVAL
    assert lib.tpt_run(None, 0, 1) == L.TPT_EINVAL
    assert lib.tpt_create(None, None) == L.TPT_EINVAL
    assert lib.tpt_batch_create(0, None, None) == L.TPT_EINVAL
    assert lib.tpt_last_error().decode()
