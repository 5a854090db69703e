"""Element tables against the reference's own known-answer tests.

Pins: crates/tpt-core/src/elements/mod.rs:1464-1509 (displacement_rules_match_tpt),
elements/mod.rs:1511-1520 (gpu_tables_have_fixed_stride), constants.rs:831-850
(PT_NUM == 256), and the weight-rank rule of move_kernel/mod.rs:1537-1550
recomputed by brute force.
"""
import numpy as np
import pytest

from tpt_hip import elements as E

P = E.pt

CAN_MOVE_GOLDEN = [
    ("NONE", "DUST", 0), ("DUST", "NONE", 1), ("SAND", "WATR", 1), ("WATR", "SAND", 0),
    ("DUST", "DUST", 0), ("DUST", "GEL", 0), ("PHOT", "NONE", 1), ("PHOT", "GLAS", 2),
    ("PHOT", "DUST", 0), ("PHOT", "GOLD", 0), ("PHOT", "BMTL", 0), ("PHOT", "LCRY", 3),
    ("PHOT", "GPMP", 3), ("NEUT", "GLAS", 2), ("NEUT", "GOLD", 2), ("ELEC", "PHOT", 2),
    ("ELEC", "VIBR", 1), ("DUST", "VIBR", 0), ("DUST", "VOID", 3), ("DUST", "PVOD", 3),
    ("DUST", "INVIS", 3), ("TRON", "SWCH", 3), ("NEUT", "INVIS", 2), ("DUST", "CNCT", 0),
    ("DUST", "BHOL", 1), ("DUST", "NBHL", 1), ("DUST", "SAWD", 0), ("EMBR", "DUST", 0),
    ("EMBR", "EMBR", 2), ("PROT", "DUST", 2), ("GRVT", "VOID", 3), ("GRVT", "INSL", 0),
    ("ANAR", "WHOL", 1), ("SPRK", "NONE", 0), ("STKM", "WATR", 2),
]


def test_pt_num_and_builtin_count():
    assert E.PT_NUM == 256
    assert sum(1 for e in E.ELEMENTS if e["enabled"]) == 239  # ids 1..239; slot 0 is NONE
    assert [e["id"] for e in E.ELEMENTS] == list(range(256))


@pytest.mark.parametrize("a,b,want", CAN_MOVE_GOLDEN)
def test_can_move_golden(a, b, want):
    assert E.build_can_move()[P(a), P(b)] == want


def test_gpu_props_known_values():
    props = E.element_props_table().reshape(256, 32)
    assert props[P("DUST"), 2] == 85          # weight
    assert props[P("WATR"), 15] == 29         # heat_conduct
    assert props.shape == (256, 32)


def test_config_element_fields():
    # Values as written in elements/data/{powders,liquids,gases,solids,special}.rs.
    d = {e["identifier"]: e for e in E.ELEMENTS}
    dust = d["DEFAULT_PT_DUST"]
    assert (dust["weight"], dust["falldown"], dust["heat_conduct"]) == (85, 1, 70)
    assert (dust["advection"], dust["air_drag"], dust["air_loss"], dust["loss"], dust["gravity"]) == \
        (0.7, 0.02, 0.96, 0.8, 0.1)
    watr = d["DEFAULT_PT_WATR"]
    assert watr["high_temp_transition"] == P("WTRV") and watr["low_temp_transition"] == P("ICEI")
    metl = d["DEFAULT_PT_METL"]
    assert metl["properties"] & E.TYPE_SOLID and metl["weight"] == 100
    wtrv = d["DEFAULT_PT_WTRV"]
    assert wtrv["properties"] & E.TYPE_GAS and wtrv["low_temp_transition"] == 512  # ST


def test_weight_rank_brute_force():
    ranks = E.weight_ranks()
    w = [e["weight"] for e in E.ELEMENTS]
    # Brute force: position = number of slots that sort strictly before.
    for i in range(256):
        pos = sum(1 for j in range(256) if (w[j], j) < (w[i], i))
        assert ranks[i] == (pos * 511) // 255
    # Heavier strictly outranks lighter; metal > dust > water > steam.
    assert ranks[P("METL")] > ranks[P("DUST")] > ranks[P("WATR")] > ranks[P("WTRV")]
    assert ranks.max() == 511 and ranks.min() == 0


def test_move_lut_layout():
    lut = E.move_lut()
    assert lut.size == 256 * 12 + 256 * 256
    b = P("DUST") * 12
    assert lut[b] == E.ELEMENTS[P("DUST")]["properties"]
    assert lut[b + 1] == 1 and lut[b + 2] == E.weight_ranks()[P("DUST")]
    assert np.array([lut[b + 4]], dtype=np.uint32).view(np.float32)[0] == np.float32(0.7)
    assert lut[256 * 12 + P("DUST") * 256 + P("NONE")] == 1
