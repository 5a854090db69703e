"""Per-array SHA-256 digests of long oracle runs of the 1600x900 and
2048x2048 workloads (run here with the C oracle, checked on the GPU).

The oracle needs minutes for what the HIP tick does in under a second, so
the GPU test compares against these digests instead of a live oracle run.
Each entry records the digest of the initial state too, so a change to the
scene generator shows up as a stale fixture rather than a parity failure.

Usage:  python tests/golden/long_run_digests.py   (writes long_run_digests.json)
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "the-powder-toy-rust-wgpu_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from tpt_hip import scene as S  # noqa: E402

FIXTURE = os.path.join(HERE, "long_run_digests.json")
# (workload, checkpoints): the tick counts of the two named workloads.
RUNS = [("1600x900", (250, 500, 1000)), ("2048x2048", (250, 500))]
SLOT = ["part_ty", "part_x", "part_y", "part_vx", "part_vy", "part_temp", "part_life", "part_ctype",
        "part_tmp", "part_tmp2", "part_flags"]
PIXEL = ["pmap", "photons", "pmap_count"]
CELL = ["air_pv", "air_vx", "air_vy", "air_hv", "wall_emap", "block_air", "block_airh", "grav_mass_prev"]


def digest(st):
    """{array: sha256 hex} over the slot range, the maps and the cell grids.
    Float zeros are normalised (+0), the only bit pattern the two sides may
    legitimately disagree on (the sign of a zero from a zero-mass solve)."""
    out = {"slot_bound": int(st.slot_bound)}
    for name in SLOT + PIXEL + CELL:
        a = st.arrays[name]
        if name in SLOT:
            a = a[: st.slot_bound]
        if a.dtype == np.float32:
            a = a + np.float32(0.0)
        out[name] = hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    return out


def main():
    from oracle import oracle as O
    result = {}
    for workload, checkpoints in RUNS:
        st = S.make_scene(workload)
        entry = {"initial": digest(st), "ticks": {}}
        o = O.OracleState(st)
        t0 = time.time()
        for frame in range(max(checkpoints)):
            o.call("oracle_step", frame)
            if frame + 1 in checkpoints:
                entry["ticks"][str(frame + 1)] = digest(st)
                print(f"{workload}: tick {frame + 1} ({time.time() - t0:.0f} s)", flush=True)
        result[workload] = entry
    with open(FIXTURE, "w") as f:
        json.dump(result, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
