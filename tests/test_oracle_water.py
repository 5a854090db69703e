"""Oracle: water equalisation (water_equal_gpu/, lib.rs:418-420).

Pinned by what TPT's flood_water fixes (Simulation.cpp:700-760): liquid only
moves down into an empty pixel on the row above the lowest free surface of
its connected body, one particle per pixel, so the particle count and the
one-particle-per-pixel map are conserved and the liquid's potential (the sum
of the rows it occupies) only grows; a communicating vessel levels out."""
import numpy as np

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import scene as S


def utube(W=64, H=64, left_top=20, right_top=45, cap=20000):
    """METL container: a left arm (x 10..20) filled from left_top, a right
    arm (x 40..50) filled from right_top, joined by a channel on rows 50..57."""
    st = S.SimState(S.default_config(W, H, capacity=cap))
    st.cfg["water_equal_test"] = 1
    a = st.arrays
    parts = []
    for y in range(4, H):
        for x in range(4, W - 4):
            wall = x < 8 or x > 55 or y > 57 or (21 <= x <= 39 and y <= 49)
            water = ((10 <= x <= 20 and y >= left_top) or (10 <= x <= 50 and 50 <= y) or
                     (40 <= x <= 50 and y >= right_top)) and y <= 57
            if water and not wall:
                parts.append(("WATR", x, y))
            elif wall and y < H - 4:
                parts.append(("METL", x, y))
    for k, (n, x, y) in enumerate(parts):
        a["part_ty"][k], a["part_x"][k], a["part_y"][k], a["part_temp"][k] = E.pt(n), x, y, 300.0
    st.slot_bound = len(parts)
    S.rebuild_pmap(st)
    return st


def liquid_rows(st):
    a = st.arrays
    ty = a["part_ty"][: st.slot_bound]
    return a["part_y"][: st.slot_bound][ty == E.pt("WATR")]


def check_map(st):
    a = st.arrays
    live = np.flatnonzero(a["part_ty"][: st.slot_bound])
    idx = a["part_y"][live].astype(np.int64) * st.cfg["xres"] + a["part_x"][live].astype(np.int64)
    assert len(np.unique(idx)) == len(live)
    assert np.array_equal((a["pmap"][idx] >> 9) - 1, live)
    assert np.count_nonzero(a["pmap"]) == len(live)


def test_utube_levels_out_and_conserves():
    st = utube()
    o = O.OracleState(st)
    rows0 = liquid_rows(st)
    potential, moved = rows0.sum(), 0
    for f in range(400):
        before = st.arrays["part_y"].copy()
        o.call("oracle_water_equal", (f * 2654435761) & 0xFFFFFFFF)
        changed = np.flatnonzero(before != st.arrays["part_y"])
        moved += len(changed)
        assert np.all(st.arrays["part_y"][changed] > before[changed])         # only downward
        p = liquid_rows(st).sum()
        assert p >= potential
        potential = p
    assert len(liquid_rows(st)) == len(rows0)
    check_map(st)
    assert moved > 50
    # The right arm's surface rose by what the left arm gave up.
    pm = st.arrays["pmap"]
    top = lambda x: min(y for y in range(64) if (pm[y * 64 + x] & 0x1FF) == E.pt("WATR"))
    assert top(45) < 45


def test_level_vessel_does_not_move():
    st = utube(left_top=45, right_top=45)
    o = O.OracleState(st)
    before = st.arrays["part_y"].copy()
    for f in range(300):
        o.call("oracle_water_equal", f * 7919)
    assert np.array_equal(before, st.arrays["part_y"])


def test_moves_land_on_the_row_above_the_lowest_free_surface():
    st = utube()
    o = O.OracleState(st)
    for f in range(200):
        before = st.arrays["part_y"].copy()
        o.call("oracle_water_equal", (f * 40503) & 0xFFFFFFFF)
        changed = np.flatnonzero(before != st.arrays["part_y"])
        for i in changed:
            # The right arm's free surface: the first moves go to row 44, x 40..50.
            assert 40 <= st.arrays["part_x"][i] <= 50
            assert st.arrays["part_y"][i] <= 44
        if len(changed):
            return
    raise AssertionError("no move in 200 frames")


def test_column_target_from_two_rows_keeps_the_highest_row():
    """Two channels join the left arm to two right arms with different
    surfaces: rows of the upper channel and rows of the lower channel assign
    the left arm's columns different targets; the lower channel (the higher
    row index) wins, so every move lands in the arm it connects to (wide
    enough that its surface stays below the other arm's meanwhile)."""
    W, H = 96, 80
    st = S.SimState(S.default_config(W, H, capacity=30000))
    st.cfg["water_equal_test"] = 1
    a = st.arrays
    parts = []
    for y in range(4, H - 4):
        for x in range(4, W - 4):
            left = 10 <= x <= 18 and 12 <= y <= 70
            upper = 19 <= x <= 50 and 40 <= y <= 44          # upper channel to arm A (x 45..50)
            arm_a = 45 <= x <= 50 and 35 <= y <= 44
            lower = 19 <= x <= 90 and 66 <= y <= 70          # lower channel to arm B (x 60..90)
            arm_b = 60 <= x <= 90 and 60 <= y <= 70
            air = (10 <= x <= 18 and y < 12) or (45 <= x <= 50 and y < 35) or (60 <= x <= 90 and y < 60)
            if left or upper or arm_a or lower or arm_b:
                parts.append(("WATR", x, y))
            elif not air:
                parts.append(("METL", x, y))
    for k, (n, x, y) in enumerate(parts):
        a["part_ty"][k], a["part_x"][k], a["part_y"][k], a["part_temp"][k] = E.pt(n), x, y, 300.0
    st.slot_bound = len(parts)
    S.rebuild_pmap(st)
    o = O.OracleState(st)
    landed = []
    for f in range(300):
        before_x, before_y = st.arrays["part_x"].copy(), st.arrays["part_y"].copy()
        o.call("oracle_water_equal", (f * 2246822519) & 0xFFFFFFFF)
        for i in np.flatnonzero(before_y != st.arrays["part_y"]):
            if 10 <= before_x[i] <= 18:
                landed.append(int(st.arrays["part_x"][i]))
    assert landed and all(60 <= x <= 90 for x in landed)
    check_map(st)
