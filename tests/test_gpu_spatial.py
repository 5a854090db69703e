"""HIP RecalcFreeParticles/CheckStacking vs the oracle (bit-exact)."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import elements as E
from tpt_hip import lib as L
from tpt_hip import scene as S

from edge_scenes import adversarial

pytestmark = pytest.mark.gpu
FIELDS = ["pmap", "photons", "pmap_count", "part_ty", "part_life", "part_tmp", "part_temp"]


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


def test_recalc_shared_pixels_with_flags_and_photons(gpu):
    """Pixels shared by several matter particles and by several photons,
    some of each at x - 0.4 (same rounded pixel, truncated pixel one to the
    left, so their keys carry the off-map flag): every full map build must
    keep the highest slot and count and list the rest as the oracle does."""
    from edge_scenes import _add
    st = S.make_scene("612x384")
    rng = np.random.default_rng(11)
    W, H = st.cfg["xres"], st.cfg["yres"]
    add = []
    for _ in range(40):
        x, y = int(rng.integers(8, W - 8)), int(rng.integers(8, H - 8))
        name = "PHOT" if rng.integers(0, 2) else str(rng.choice(["DUST", "WATR", "METL"]))
        for _ in range(int(rng.integers(2, 9))):
            add.append((name, float(x) - (0.4 if rng.integers(0, 2) else 0.0), y, {"life": 0}))
    _add(st, add)
    ref = st.copy()
    o = O.OracleState(ref)
    print("[recalc parity] shared pixels", flush=True)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for frame in range(6):
            seed = L.frame_seed(frame)
            ctx.recalc_free_particles(seed, 0)
            o.call("oracle_recalc_free_particles", seed, 0)
        ctx.fetch(st, FIELDS)
    for n in FIELDS:
        a, b = ref.arrays[n].view(np.uint32), st.arrays[n].view(np.uint32)
        assert np.array_equal(a, b), f"{n}: {np.count_nonzero(a != b)} differ"
