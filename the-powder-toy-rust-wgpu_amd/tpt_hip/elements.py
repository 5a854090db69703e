"""Element registry host mirror: element rows, ``can_move`` and ``move_lut``.

The reference host builds these tables in Rust and uploads them:

* ``GpuElementProps`` rows, 32 words per element
  (``crates/tpt-core/src/elements/mod.rs:482-556``);
* the ``can_move`` displacement matrix
  (``crates/tpt-core/src/elements/mod.rs:719-866``, ``rebuild_can_move``);
* the packed movement LUT with the weight rank
  (``crates/tpt-sim/src/move_kernel/mod.rs:1534-1576``, ``build_move_lut``).

The element data itself comes from ``data/elements.json`` (text-extracted from
``crates/tpt-core/src/elements/data/*.rs`` by ``data/gen_elements.py``).
"""

import json
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "elements.json")

ELEM_STRIDE = 32          # words per GpuElementProps row
META_WORDS = 12           # words per element record in move_lut
RANK_MASK = 0x1FF


def _load():
    with open(_DATA) as f:
        return json.load(f)


_J = _load()
PT_NUM = int(_J["pt_num"])
ELEMENTS = _J["elements"]
C = _J["constants"]          # PT_*, TYPE_*, PROP_*, FLAG_*, WL_* as parsed

TYPE_PART, TYPE_LIQUID, TYPE_SOLID, TYPE_GAS, TYPE_ENERGY = 0x1, 0x2, 0x4, 0x8, 0x10


def pt(name):
    """Element id by its short name, e.g. ``pt("DUST") == 1``."""
    return int(C["PT_" + name])


def _f32_bits(v):
    return int(np.array([v], dtype=np.float32).view(np.uint32)[0])


def _i32_bits(v):
    return int(np.array([v], dtype=np.int32).view(np.uint32)[0])


def element_props_table(elements=None):
    """``GpuElementProps::from_def`` for all slots -> uint32[PT_NUM*32]."""
    elements = ELEMENTS if elements is None else elements
    out = np.zeros(PT_NUM * ELEM_STRIDE, dtype=np.uint32)
    for e in elements:
        b = e["id"] * ELEM_STRIDE
        row = [
            e["colour"] & 0xFFFFFFFF, e["properties"], _i32_bits(e["weight"]), e["falldown"],
            _f32_bits(e["advection"]), _f32_bits(e["air_drag"]), _f32_bits(e["air_loss"]),
            _f32_bits(e["loss"]), _f32_bits(e["collision"]), _f32_bits(e["gravity"]),
            _f32_bits(e["diffusion"]), _f32_bits(e["hot_air"]),
            _i32_bits(e["flammable"]), _i32_bits(e["explosive"]), _i32_bits(e["hardness"]),
            e["heat_conduct"], _f32_bits(e["default_temp"]), _i32_bits(e["default_life"]),
            _f32_bits(e["low_pressure"]), _i32_bits(e["low_pressure_transition"]),
            _f32_bits(e["high_pressure"]), _i32_bits(e["high_pressure_transition"]),
            _f32_bits(e["low_temp"]), _i32_bits(e["low_temp_transition"]),
            _f32_bits(e["high_temp"]), _i32_bits(e["high_temp_transition"]),
            _i32_bits(e["meltable"]), e["default_ctype"] & 0xFFFFFFFF,
            e["pixel_mode"], e["fire_colour"] & 0xFFFFFFFF, 255, e["carries_ctype"],
        ]
        out[b:b + ELEM_STRIDE] = row
    return out


def build_can_move(elements=None):
    """``ElementRegistry::rebuild_can_move`` (elements/mod.rs:719-866) -> uint8[PT_NUM, PT_NUM]."""
    E = ELEMENTS if elements is None else elements
    n = PT_NUM
    t = np.zeros((n, n), dtype=np.uint8)
    props = [e["properties"] for e in E]
    weight = [e["weight"] for e in E]
    P = pt
    t[1:, :] = 1
    t[P("PHOT"), 1:] = 2
    for m in range(1, n):
        for d in range(1, n):
            if weight[m] <= weight[d] or d == P("GEL"):
                t[m, d] = 0
            if m == P("NEUT"):
                if props[d] & C["PROP_NEUTPASS"]:
                    t[m, d] = 2
                if props[d] & C["PROP_NEUTABSORB"]:
                    t[m, d] = 1
                if props[d] & C["PROP_NEUTPENETRATE"]:
                    t[m, d] = 1
            if d == P("NEUT") and props[m] & C["PROP_NEUTPENETRATE"]:
                t[m, d] = 0
            if props[m] & TYPE_ENERGY and props[d] & TYPE_ENERGY:
                t[m, d] = 2
    for d in range(n):
        stkm = 0
        if props[d] & (TYPE_LIQUID | TYPE_GAS):
            stkm = 2
        if d == 0 or d in (P("PRTO"), P("SPAWN"), P("SPAWN2")):
            stkm = 2
        t[P("STKM"), d] = stkm
        t[P("STKM2"), d] = stkm
        t[P("FIGH"), d] = stkm
        t[P("SPRK"), d] = 0
    for m in range(1, n):
        t[m, P("BHOL")] = 1
        t[m, P("NBHL")] = 1
        t[m, P("STKM")] = 0
        t[m, P("STKM2")] = 0
        t[m, P("FIGH")] = 0
        t[m, P("INVIS")] = 3
        t[m, P("CNCT")] = 0
        t[m, P("PVOD")] = 3
        t[m, P("VOID")] = 3
        t[m, P("EMBR")] = 0
        t[P("EMBR"), m] = 0
        if props[m] & TYPE_ENERGY:
            t[m, P("VIBR")] = 1
            t[m, P("BVBR")] = 1
        if props[m] & TYPE_PART:
            t[m, P("SAWD")] = 0
    blockers = {P(x) for x in ("DMND", "INSL", "VOID", "PVOD", "VIBR", "BVBR", "PRTI", "PRTO")}
    for d in range(n):
        if props[d] & C["PROP_PHOTPASS"]:
            t[P("PHOT"), d] = 2
        if d not in blockers:
            t[P("PROT"), d] = 2
            t[P("GRVT"), d] = 2
    for d in ("DMND", "CLNE", "PCLN", "BCLN", "PBCN", "ROCK"):
        t[P("DEST"), P(d)] = 0
    pairs = [
        ("NEUT", "INVIS", 2), ("ELEC", "LCRY", 2), ("ELEC", "EXOT", 2), ("ELEC", "GLOW", 2),
        ("PHOT", "LCRY", 3), ("PHOT", "GPMP", 3), ("PHOT", "BIZR", 2), ("ELEC", "BIZR", 2),
        ("PHOT", "BIZRG", 2), ("ELEC", "BIZRG", 2), ("PHOT", "BIZRS", 2), ("ELEC", "BIZRS", 2),
        ("BIZR", "FILT", 2), ("BIZRG", "FILT", 2), ("ANAR", "WHOL", 1), ("ANAR", "NWHL", 1),
        ("ELEC", "DEUT", 1), ("THDR", "THDR", 2), ("EMBR", "EMBR", 2), ("TRON", "SWCH", 3),
        ("ELEC", "RSST", 2), ("ELEC", "RSSS", 2),
    ]
    for a, b, v in pairs:
        t[P(a), P(b)] = v
    return t


def weight_ranks(elements=None):
    """9-bit weight rank: position in the ``(weight, id)`` sort scaled to 0..511
    (move_kernel/mod.rs:1537-1550)."""
    E = ELEMENTS if elements is None else elements
    order = sorted(range(PT_NUM), key=lambda i: (E[i]["weight"], i))
    last = max(PT_NUM - 1, 1)
    ranks = np.zeros(PT_NUM, dtype=np.uint32)
    for pos, i in enumerate(order):
        ranks[i] = ((pos * RANK_MASK) // last) & RANK_MASK
    return ranks


def move_lut(elements=None):
    """``build_move_lut`` -> uint32[PT_NUM*12 + PT_NUM*PT_NUM]."""
    E = ELEMENTS if elements is None else elements
    lut = np.zeros(PT_NUM * META_WORDS + PT_NUM * PT_NUM, dtype=np.uint32)
    ranks = weight_ranks(E)
    for e in E:
        b = e["id"] * META_WORDS
        lut[b:b + META_WORDS] = [
            e["properties"], e["falldown"], ranks[e["id"]], e["heat_conduct"],
            _f32_bits(e["advection"]), _f32_bits(e["air_drag"]), _f32_bits(e["air_loss"]),
            _f32_bits(e["loss"]), _f32_bits(e["collision"]), _f32_bits(e["gravity"]),
            _f32_bits(e["diffusion"]), _f32_bits(e["hot_air"]),
        ]
    lut[PT_NUM * META_WORDS:] = build_can_move(E).astype(np.uint32).reshape(-1)
    return lut


def photon_mask(elements=None):
    """``build_photon_mask`` (move_kernel/mod.rs:1592-1606) -> uint32[PT_NUM]."""
    E = ELEMENTS if elements is None else elements
    return np.array([e["photon_reflect"] & 0xFFFFFFFF for e in E], dtype=np.uint32)


SPAWN_DEFAULT_WORDS = 8


def spawn_defaults(elements=None):
    """``to_gpu_spawn_defaults_u32`` (tpt-core/src/elements/mod.rs:900-936) ->
    uint32[PT_NUM * 8]: {temp bits, life, ctype, tmp, tmp2, tmp3, tmp4, 0}."""
    E = ELEMENTS if elements is None else elements
    out = np.zeros(PT_NUM * SPAWN_DEFAULT_WORDS, dtype=np.uint32)
    for t, e in enumerate(E):
        out[t * SPAWN_DEFAULT_WORDS: (t + 1) * SPAWN_DEFAULT_WORDS] = [
            _f32_bits(e["default_temp"]), _i32_bits(e["default_life"]), e["default_ctype"] & 0xFFFFFFFF,
            _i32_bits(e["default_tmp"]), _i32_bits(e["default_tmp2"]), _i32_bits(e["default_tmp3"]),
            _i32_bits(e["default_tmp4"]), 0]
    return out


def _cached_default(fn):
    """The tables of the built-in registry are constant: build them once and
    hand out copies (callers may edit theirs, e.g. a LUT with zeroed ranks)."""
    cache = {}

    def wrapper(elements=None):
        if elements is not None:
            return fn(elements)
        if "v" not in cache:
            cache["v"] = fn(None)
        return cache["v"].copy()

    wrapper.__doc__ = fn.__doc__
    wrapper.__name__ = fn.__name__
    return wrapper


element_props_table = _cached_default(element_props_table)
build_can_move = _cached_default(build_can_move)
weight_ranks = _cached_default(weight_ranks)
move_lut = _cached_default(move_lut)
photon_mask = _cached_default(photon_mask)
spawn_defaults = _cached_default(spawn_defaults)
