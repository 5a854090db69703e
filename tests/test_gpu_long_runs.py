"""The 1600x900 workload for its full 1000 ticks and the 2048x2048 workload
for its full 500 ticks, through graph replays (tpt_run), against per-array
SHA-256 digests of the oracle's runs at ticks 250/500(/1000), written by
tests/golden/long_run_digests.py."""
import json

import pytest

from golden.long_run_digests import FIXTURE, digest
from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("workload", ["1600x900", "2048x2048"])
def test_long_run_matches_oracle_digests(gpu, workload):
    with open(FIXTURE) as f:
        fx = json.load(f)[workload]
    st = S.make_scene(workload)
    assert digest(st) == fx["initial"], "the scene generator changed: regenerate tests/golden/long_run_digests.json"
    print(f"[long run] {workload}", flush=True)
    done = 0
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        for t in sorted(int(k) for k in fx["ticks"]):
            ctx.run(done, t - done)
            done = t
            ctx.fetch(st)
            got, want = digest(st), fx["ticks"][str(t)]
            bad = sorted(k for k in want if got.get(k) != want[k])
            print(f"  tick {t}: {len(want) - len(bad)}/{len(want)} arrays match", flush=True)
            assert not bad, f"{workload} tick {t}: {bad} differ from the oracle"
