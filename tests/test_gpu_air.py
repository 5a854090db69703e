"""HIP air step vs the CPU oracle on the same inputs (bit-exact).

The HIP kernels (csrc/air.hip) fuse the reference's seven air launches
(air_kernel/mod.rs:150-291) into four; every field must come out identical to
the oracle's kernel-by-kernel restatement, tick after tick."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import lib as L
from tpt_hip import scene as S

AIR_FIELDS = ["air_vx", "air_vy", "air_pv", "air_hv", "block_air", "block_airh", "wall_emap"]

pytestmark = pytest.mark.gpu


def _compare(st_ref, st_gpu, names, label):
    for n in names:
        a, b = st_ref.arrays[n], st_gpu.arrays[n]
        if not np.array_equal(a.view(np.uint32), b.view(np.uint32)):
            bad = np.flatnonzero(a.view(np.uint32) != b.view(np.uint32))
            i = bad[0]
            raise AssertionError(f"{label}: {n} differs at {len(bad)} cells, first {i}: oracle {a[i]!r} hip {b[i]!r}")


@pytest.mark.parametrize("workload,scale,ticks", [("612x384", 1.0, 30), ("2048x2048", 0.02, 10)])
@pytest.mark.parametrize("aheat", [True, False])
def test_air_step_bit_exact(gpu, workload, scale, ticks, aheat):
    st = S.make_scene(workload, scale_particles=scale, aheat=aheat)
    # A powered EWALL strip and a fan exercise the emap decay and fan paths.
    w = st.arrays["wall_bmap"].reshape(st.ycells, st.xcells)
    w[10:12, 20:40] = 2
    st.arrays["wall_emap"].reshape(st.ycells, st.xcells)[10:12, 20:30] = 3
    w[30, 30:40] = 5
    st.arrays["fan_vx"].reshape(st.ycells, st.xcells)[30, 30:40] = 0.5
    print(f"[air parity] {workload} aheat={aheat} ticks={ticks}", flush=True)
    ref = st.copy()
    o = O.OracleState(ref)
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for t in range(ticks):
            o.call("oracle_step_air")
            ctx.step_air()
            if t in (0, ticks - 1):
                ctx.fetch(st, AIR_FIELDS)
                _compare(ref, st, AIR_FIELDS, f"{workload} tick {t}")


def test_air_mode_variants(gpu):
    for mode in (1, 2, 3, 4):
        st = S.make_scene("612x384", scale_particles=0.2)
        st.cfg["air_mode"] = mode
        ref = st.copy()
        o = O.OracleState(ref)
        with L.Context(st.cfg) as ctx:
            ctx.load_state(st)
            for _ in range(3):
                o.call("oracle_step_air")
                ctx.step_air()
            ctx.fetch(st, AIR_FIELDS)
        _compare(ref, st, AIR_FIELDS, f"air_mode {mode}")
