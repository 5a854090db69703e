"""HIP RecalcFreeParticles/CheckStacking vs the oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu
FIELDS = ["pmap", "photons", "pmap_count", "part_ty", "part_life", "part_tmp", "part_temp"]


def adversarial(st, rng):
    """Stacks, INVIS/FILT pairs, photons and FIRE with short life on top of a scene."""
    a = st.arrays
    free = np.flatnonzero(a["part_ty"] == 0)
    base = st.slot_bound
    add = []
    for k in range(20):                      # stacks of 3..60 matter particles
        x, y = int(rng.integers(8, st.cfg["xres"] - 8)), int(rng.integers(8, st.cfg["yres"] - 8))
        for _ in range(int(rng.integers(3, 60))):
            add.append(("DUST", x, y, 0))
    for k in range(20):                      # INVIS/FILT sharing pixels with matter or each other
        x, y = int(rng.integers(8, st.cfg["xres"] - 8)), int(rng.integers(8, st.cfg["yres"] - 8))
        for name in rng.choice(["INVIS", "FILT", "WATR"], size=3):
            add.append((str(name), x, y, 0))
    for k in range(50):
        add.append(("PHOT", int(rng.integers(8, 200)), int(rng.integers(8, 200)), 0))
        add.append(("FIRE", int(rng.integers(8, 200)), int(rng.integers(8, 200)), int(rng.integers(0, 3))))
    for j, (name, x, y, life) in enumerate(add):
        s = base + j
        a["part_ty"][s], a["part_x"][s], a["part_y"][s], a["part_life"][s] = E.pt(name), x, y, life
    st.slot_bound = base + len(add)


@pytest.mark.parametrize("workload,scale", [("612x384", 1.0), ("2048x2048", 0.25)])
def test_recalc_bit_exact(gpu, workload, scale):
    st = S.make_scene(workload, scale_particles=scale)
    adversarial(st, np.random.default_rng(5))
    ref = st.copy()
    o = O.OracleState(ref)
    print(f"[recalc parity] {workload}", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(12):
            seed = L.frame_seed(frame)
            ctx.recalc_free_particles(seed, frame % 10 == 0)
            o.call("oracle_recalc_free_particles", seed, 1 if frame % 10 == 0 else 0)
        ctx.fetch(st, FIELDS)
    for n in FIELDS:
        a, b = ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)
        assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} differ"
