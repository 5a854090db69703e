"""Scene edits that reach the rare branches of the map and thermal passes:
stacked pixels, INVIS/FILT, energy particles, short-lived FIRE, the element
types whose conduction reads life/tmp/ctype, and off-pixel positions."""
import numpy as np

from tpt_hip import elements as E


def _add(st, items):
    a = st.arrays
    base = st.slot_bound
    for j, it in enumerate(items):
        name, x, y = it[:3]
        extra = it[3] if len(it) > 3 else {}
        s = base + j
        a["part_ty"][s], a["part_x"][s], a["part_y"][s] = E.pt(name), x, y
        a["part_temp"][s] = extra.get("temp", 300.0)
        for k in ("life", "tmp", "ctype"):
            if k in extra:
                a["part_" + k][s] = extra[k]
    st.slot_bound = base + len(items)


def adversarial(st, rng):
    """Stacks, INVIS/FILT pairs, photons and FIRE with short life."""
    W, H = st.cfg["xres"], st.cfg["yres"]
    add = []
    for _ in range(20):                      # stacks of 3..60 matter particles
        x, y = int(rng.integers(8, W - 8)), int(rng.integers(8, H - 8))
        for _ in range(int(rng.integers(3, 60))):
            add.append(("DUST", x, y, {"life": 0}))
    for _ in range(20):                      # INVIS/FILT sharing pixels with matter or each other
        x, y = int(rng.integers(8, W - 8)), int(rng.integers(8, H - 8))
        for name in rng.choice(["INVIS", "FILT", "WATR"], size=3):
            add.append((str(name), x, y, {"life": 0}))
    for _ in range(50):
        add.append(("PHOT", int(rng.integers(8, 200)), int(rng.integers(8, 200)), {"life": 0}))
        add.append(("FIRE", int(rng.integers(8, 200)), int(rng.integers(8, 200)),
                    {"life": int(rng.integers(0, 3))}))
    _add(st, add)


def thermal_specials(st, rng):
    """3x3 clusters mixing HSWC/PIPE/PPIP/GEL/BASE/FILT/BRAY/ELEC/DEUT with
    ordinary conductors, at varied life/tmp/ctype and temperatures."""
    W, H = st.cfg["xres"], st.cfg["yres"]
    kinds = ["HSWC", "PIPE", "PPIP", "GEL", "BASE", "FILT", "BRAY", "ELEC", "DEUT", "METL", "WATR", "DUST"]
    add = []
    for _ in range(60):
        cx, cy = int(rng.integers(10, W - 10)), int(rng.integers(10, H - 10))
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                name = str(rng.choice(kinds))
                extra = {"temp": float(rng.uniform(200.0, 2000.0)),
                         "life": int(rng.choice([0, 10, 3])),
                         "tmp": int(rng.choice([0, 1, 2, 3, 40, 100])),
                         "ctype": int(rng.choice([0, E.pt("BASE"), E.pt("WATR")]))}
                add.append((name, cx + dx, cy + dy, extra))
    _add(st, add)


def off_pixel_positions(st, rng, n=400):
    """Moves some live particles to fractional positions: rounded and
    truncated pixels equal (+0.3), different (+0.6), and at the canvas edges
    (x = -1.0 rounds onto the canvas, y = yres - 0.4 truncates onto it)."""
    a = st.arrays
    W, H = st.cfg["xres"], st.cfg["yres"]
    live = np.flatnonzero(a["part_ty"][: st.slot_bound])
    pick = rng.choice(live, size=min(n, len(live)), replace=False)
    q = len(pick) // 4
    a["part_x"][pick[:q]] += np.float32(0.3)
    a["part_y"][pick[q:2 * q]] += np.float32(0.6)
    a["part_x"][pick[2 * q:3 * q]] = np.float32(-1.0)
    a["part_y"][pick[3 * q:]] = np.float32(H - 0.4)
