"""Host side of the learned air surrogate (tpt_sur_* in include/tpt_hip.h).

The network advances a context's air grids (pressure, velocity, ambient heat)
by 8 plain air ticks per call.  This module builds its training set with the
plain air step (tpt_step_air, air_kernel::step_air_gpu in the reference,
air_kernel/mod.rs:120-299), drives training on the GPU, evaluates the trained
network against the plain solver, and saves and loads its parameters.

    python -m tpt_hip.surrogate train --traj 2000 --epochs 6 --out weights.npz
    python -m tpt_hip.surrogate train --init weights.npz --rollout 3 --epochs 4 --out tuned.npz
    python -m tpt_hip.surrogate eval --weights weights.npz --seeds 256
    python -m tpt_hip.surrogate calls --calls 30      (a workload for rocprofv3)

Scenes: 4-16 Gaussian pressure blobs (amplitude +-30, sigma 4-10 cells),
0-32 WL_WALL rectangles (sides 2-32 cells) inside a border ring, 0-8 heat
sources (squares of 1-4 cells held at 500-3000 K).  All air grids start at
zero velocity and ambient heat.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import lib as L
from . import scene as S

TICKS_PER_CALL = 8
# input normalisation: pressure, vx, vy, heat above ambient (the scenes'
# pressure blobs reach +-30, velocities a few cells per tick, sources 3000 K)
IN_SCALES = (10.0, 1.0, 1.0, 500.0)
MAX_BLOBS, MAX_RECTS, MAX_SOURCES = 16, 32, 8
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "air_surrogate_512.npz")


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_blobs", ctypes.c_uint32), ("n_rects", ctypes.c_uint32), ("n_sources", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("blob", (ctypes.c_float * 4) * MAX_BLOBS),
        ("rect", (ctypes.c_uint32 * 4) * MAX_RECTS),
        ("source", (ctypes.c_uint32 * 4) * MAX_SOURCES),
        ("source_temp", ctypes.c_float * MAX_SOURCES),
    ]


def random_scene(seed, xc, yc):
    """The scene of `seed` on an xc x yc air grid."""
    rng = np.random.default_rng(seed)
    d = SceneDesc()
    d.n_blobs = int(rng.integers(4, 17))
    for b in range(d.n_blobs):
        d.blob[b][0], d.blob[b][1] = float(rng.uniform(0, xc)), float(rng.uniform(0, yc))
        d.blob[b][2], d.blob[b][3] = float(rng.uniform(-30, 30)), float(rng.uniform(4, 10))
    d.n_rects = int(rng.integers(0, 33))
    for r in range(d.n_rects):
        w, h = int(rng.integers(2, 33)), int(rng.integers(2, 33))
        x0, y0 = int(rng.integers(1, max(2, xc - w - 1))), int(rng.integers(1, max(2, yc - h - 1)))
        d.rect[r][:] = [x0, y0, x0 + w, y0 + h]
    d.n_sources = int(rng.integers(0, 9))
    for k in range(d.n_sources):
        s = int(rng.integers(1, 5))
        x0, y0 = int(rng.integers(3, xc - s - 3)), int(rng.integers(3, yc - s - 3))
        d.source[k][:] = [x0, y0, x0 + s, y0 + s]
        d.source_temp[k] = float(rng.uniform(500, 3000))
    return d


def air_config(xc, yc):
    """A particle-free context whose air grid is xc x yc cells, the plain
    tick's defaults (ambient heat and Boussinesq convection on)."""
    c = S.default_config(4 * xc, 4 * yc, capacity=65536)
    c["aheat_enable"] = 1
    c["newton_grav"] = 0
    return c


class Surrogate:
    """A tpt_sur network for an xc x yc air grid."""

    def __init__(self, xc, yc, in_scales=IN_SCALES, out_scales=(1.0, 1.0, 1.0, 10.0), ambient=295.15):
        self.xc, self.yc = xc, yc
        self._h = ctypes.c_void_p()
        L._check(L.load().tpt_sur_create(xc, yc, ctypes.byref(self._h)))
        n = ctypes.c_size_t()
        L._check(L.load().tpt_sur_param_count(self._h, ctypes.byref(n)))
        self.n_params = n.value
        self.set_scales(in_scales, out_scales, ambient)

    def set_scales(self, in_scales, out_scales, ambient=295.15):
        self.in_scales = tuple(float(s) for s in in_scales)
        self.out_scales = tuple(float(s) for s in out_scales)
        self.ambient = float(ambient)
        L._check(L.load().tpt_sur_set_scales(self._h, (ctypes.c_float * 4)(*self.in_scales),
                                             (ctypes.c_float * 4)(*self.out_scales), ctypes.c_float(self.ambient)))

    def close(self):
        if self._h:
            L.load().tpt_sur_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.n_params, np.float32)
        L._check(L.load().tpt_sur_get_params(self._h, out.ctypes.data, out.size))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        L._check(L.load().tpt_sur_set_params(self._h, p.ctypes.data, p.size))

    def init_params(self, seed):
        L._check(L.load().tpt_sur_init_params(self._h, seed))

    def grads(self):
        out = np.empty(self.n_params, np.float32)
        L._check(L.load().tpt_sur_get_grads(self._h, out.ctypes.data, out.size))
        return out

    def forward_raw(self, x0_bits):
        x0_bits = np.ascontiguousarray(x0_bits, dtype=np.uint16)
        out = np.empty(self.xc * self.yc * 8, np.float32)
        L._check(L.load().tpt_sur_forward_raw(self._h, x0_bits.ctypes.data, out.ctypes.data))
        return out

    def backward_raw(self, g_bits):
        g_bits = np.ascontiguousarray(g_bits, dtype=np.uint16)
        L._check(L.load().tpt_sur_backward_raw(self._h, g_bits.ctypes.data))

    def scene(self, ctx, desc):
        L._check(L.load().tpt_sur_scene(self._h, ctx._h, ctypes.byref(desc)))

    def plain_ticks(self, ctx, ticks):
        L._check(L.load().tpt_sur_plain_ticks(self._h, ctx._h, ticks))

    def dataset_alloc(self, n_traj, n_snap):
        L._check(L.load().tpt_sur_dataset_alloc(self._h, n_traj, n_snap))

    def capture(self, ctx, traj, snap):
        L._check(L.load().tpt_sur_dataset_capture(self._h, ctx._h, traj, snap))

    def train(self, pairs, accum, lr):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1)
        loss = ctypes.c_float()
        L._check(L.load().tpt_sur_train(self._h, pairs.ctypes.data, pairs.size // 2, accum, ctypes.c_float(lr),
                                        ctypes.byref(loss)))
        return loss.value

    def set_clip(self, clip):
        L._check(L.load().tpt_sur_set_clip(self._h, ctypes.c_float(clip)))

    def set_rollout(self, steps):
        """Training horizon: 1 (snapshot k -> k+1) or 2 (the network's own
        prediction of snapshot k+1 -> snapshot k+2)."""
        L._check(L.load().tpt_sur_set_rollout(self._h, steps))

    def set_fused(self, fused):
        L._check(L.load().tpt_sur_set_fused(self._h, 1 if fused else 0))

    def eval_loss(self, pairs):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1)
        loss = ctypes.c_float()
        L._check(L.load().tpt_sur_eval_loss(self._h, pairs.ctypes.data, pairs.size // 2, ctypes.byref(loss)))
        return loss.value

    def step_air(self, ctx, check=True, energy_ratio=1.05):
        fb = ctypes.c_int()
        L._check(L.load().tpt_sur_step_air(self._h, ctx._h, 1 if check else 0, ctypes.c_float(energy_ratio),
                                           ctypes.byref(fb)))
        return bool(fb.value)

    def save(self, path, meta=None):
        np.savez(path, params=self.params(), in_scales=np.float32(self.in_scales),
                 out_scales=np.float32(self.out_scales), ambient=np.float32(self.ambient),
                 grid=np.int32([self.xc, self.yc]), meta=json.dumps(meta or {}))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_scales(tuple(z["in_scales"]), tuple(z["out_scales"]), float(z["ambient"]))
            self.set_params(z["params"])
            return json.loads(str(z["meta"]))


def air_state(ctx, n):
    return {k: ctx.download(f"air_{k}", np.float32, n) for k in ("pv", "vx", "vy", "hv")}


def build_dataset(sur, ctx, seeds, n_snap=TICKS_PER_CALL + 1, log=None):
    """Trajectories of the plain air step, one snapshot every 8 ticks."""
    sur.dataset_alloc(len(seeds), n_snap)
    for t, seed in enumerate(seeds):
        sur.scene(ctx, random_scene(seed, sur.xc, sur.yc))
        for k in range(n_snap):
            if k:
                sur.plain_ticks(ctx, TICKS_PER_CALL)
            sur.capture(ctx, t, k)
        if log and (t + 1) % 250 == 0:
            ctx.synchronize()
            log(f"dataset: {t + 1}/{len(seeds)} trajectories")
    ctx.synchronize()


def delta_scales(ctx, sur, seeds, ticks=TICKS_PER_CALL):
    """Per-field standard deviation of the 8-tick change over a few scenes:
    the normalisation of the network's inputs and outputs."""
    n = sur.xc * sur.yc
    ds = {k: [] for k in ("pv", "vx", "vy", "hv")}
    for seed in seeds:
        sur.scene(ctx, random_scene(seed, sur.xc, sur.yc))
        for _ in range(4):
            a = air_state(ctx, n)
            sur.plain_ticks(ctx, ticks)
            b = air_state(ctx, n)
            for k in ds:
                ds[k].append(b[k] - a[k])
    sp = float(np.std(np.concatenate(ds["pv"])))
    sv = float(np.std(np.concatenate(ds["vx"] + ds["vy"])))
    # heat changes only near the sources: the spread of the cells that change
    h = np.concatenate(ds["hv"])
    h = h[np.abs(h) > 1e-3]
    sh = float(np.std(h)) if h.size else 1.0
    return (max(sp, 1e-3), max(sv, 1e-3), max(sv, 1e-3), max(sh, 1e-3))


FIELDS = ("pv", "vx", "vy", "hv")


def sq_errors(pred, ref, ambient):
    """Per field: (squared L2 error, squared L2 norm of the reference), heat
    measured from ambient."""
    out = {}
    for k in FIELDS:
        a, b = pred[k].astype(np.float64), ref[k].astype(np.float64)
        if k == "hv":
            a, b = a - ambient, b - ambient
        out[k] = (float(np.sum((a - b) ** 2)), float(np.sum(b * b)))
    return out


def rel_l2(sums):
    """Relative L2 error per field over a whole evaluation set (errors and
    reference norms summed over its scenes), and the mean over the fields."""
    out = {k: float(np.sqrt(sums[k][0] / sums[k][1])) if sums[k][1] > 0 else 0.0 for k in FIELDS}
    out["mean"] = float(np.mean([out[k] for k in FIELDS]))
    return out


def evaluate(sur, ctx, seeds, calls=8, check=True, energy_ratio=1.05, persistence=False):
    """Relative L2 error after `calls` surrogate calls (8 ticks each) against
    the plain solver from the same scene, per seed; fallback count."""
    n = sur.xc * sur.yc
    sums = {k: [0.0, 0.0] for k in FIELDS}
    first = {k: [0.0, 0.0] for k in FIELDS}   # after the first call (8 ticks)
    fallbacks, total = 0, 0
    for seed in seeds:
        desc = random_scene(seed, sur.xc, sur.yc)
        sur.scene(ctx, desc)
        sur.plain_ticks(ctx, TICKS_PER_CALL)
        ref1 = air_state(ctx, n)
        sur.plain_ticks(ctx, TICKS_PER_CALL * (calls - 1))
        ref = air_state(ctx, n)
        sur.scene(ctx, desc)
        for c in range(calls):
            if not persistence:
                fallbacks += sur.step_air(ctx, check=check, energy_ratio=energy_ratio)
                total += 1
            if c == 0:
                for k, (e, r) in sq_errors(air_state(ctx, n), ref1, sur.ambient).items():
                    first[k][0] += e
                    first[k][1] += r
        for k, (e, r) in sq_errors(air_state(ctx, n), ref, sur.ambient).items():
            sums[k][0] += e
            sums[k][1] += r
    return {"rel_l2": rel_l2(sums), "rel_l2_8_ticks": rel_l2(first), "fallbacks": fallbacks, "calls": total,
            "fallback_rate": (fallbacks / total) if total else 0.0, "seeds": len(seeds)}


def timing(sur, ctx, seed=12345, reps=20):
    """Mean time of one surrogate call (unchecked and checked) and of 8 plain
    air ticks on the same scene, from host timers around synchronised runs."""
    sur.scene(ctx, random_scene(seed, sur.xc, sur.yc))
    for _ in range(3):
        sur.step_air(ctx, check=False)
        sur.plain_ticks(ctx, TICKS_PER_CALL)
    ctx.synchronize()
    out = {}
    for name, fn in (("surrogate_call", lambda: sur.step_air(ctx, check=False)),
                     ("surrogate_call_checked", lambda: sur.step_air(ctx, check=True)),
                     ("plain_8_ticks", lambda: sur.plain_ticks(ctx, TICKS_PER_CALL))):
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        ctx.synchronize()
        out[name] = (time.perf_counter() - t0) / reps * 1e6
    return out


def train_main(args):
    log = lambda m: print(f"[surrogate {time.strftime('%H:%M:%S')}] {m}", flush=True)  # noqa: E731
    xc = yc = args.grid
    ctx = L.Context(air_config(xc, yc))
    ctx.load_state(S.SimState(air_config(xc, yc)))
    sur = Surrogate(xc, yc)
    history, prior = [], {}
    if args.init:
        # fine-tuning: parameters and normalisation of a saved network
        prior = sur.load(args.init)
        history = list(prior.get("history", []))
        log(f"starting from {args.init} ({prior.get('epochs')} epochs)")
    else:
        scales = delta_scales(ctx, sur, range(10 ** 6, 10 ** 6 + 8))
        sur.set_scales(IN_SCALES, scales)
        log(f"output scales {scales}")
        sur.init_params(args.seed)
    sur.set_clip(args.clip)
    train_seeds = list(range(args.traj))
    t0 = time.perf_counter()
    build_dataset(sur, ctx, train_seeds, log=log)
    log(f"dataset built in {time.perf_counter() - t0:.1f} s")
    pairs = np.array([(t, k) for t in range(args.traj) for k in range(TICKS_PER_CALL)], np.uint32)
    rng = np.random.default_rng(args.seed)
    hold = pairs[rng.choice(len(pairs), size=min(256, len(pairs)), replace=False)]
    steps_done = 0
    total_steps = args.epochs * sum(int(np.count_nonzero(pairs[:, 1] + r <= TICKS_PER_CALL))
                                    for r in range(1, args.rollout + 1)) // args.accum
    for ep in range(args.epochs):
        order = rng.permutation(len(pairs))
        t1 = time.perf_counter()
        chunk = args.accum * 64
        losses = []
        for c0 in range(0, len(order), chunk):
            frac = steps_done / max(total_steps, 1)
            warm = min(1.0, (steps_done + 1) / args.warmup_steps)
            lr = args.lr * warm * 0.5 * (1.0 + np.cos(np.pi * min(frac, 1.0)))
            sel = pairs[order[c0:c0 + chunk]]
            sur.set_rollout(1)
            losses.append(sur.train(sel, args.accum, lr))
            steps_done += (len(sel) + args.accum - 1) // args.accum
            for r in range(2, args.rollout + 1):
                # the same pairs again, from the network's own r - 1 steps
                deep = sel[sel[:, 1] + r <= TICKS_PER_CALL]
                if len(deep):
                    sur.set_rollout(r)
                    losses.append(sur.train(deep, args.accum, lr))
                    sur.set_rollout(1)
                    steps_done += (len(deep) + args.accum - 1) // args.accum
        ev = sur.eval_loss(hold)
        done_epochs = int(prior.get("epochs") or 0) + ep + 1
        history.append({"epoch": done_epochs, "train_loss": float(np.mean(losses)), "eval_loss": ev,
                        "seconds": time.perf_counter() - t1, "rollout": args.rollout})
        log(f"epoch {done_epochs}: train {np.mean(losses):.4e} eval {ev:.4e} ({time.perf_counter() - t1:.1f} s)")
        sur.save(args.out, {"history": history, "traj": args.traj, "epochs": done_epochs, "seed": args.seed})
    log(f"saved {args.out}")
    if args.eval_seeds:
        res = evaluate(sur, ctx, range(10 ** 7, 10 ** 7 + args.eval_seeds))
        log(f"held-out {args.eval_seeds} seeds: {res}")
        sur.save(args.out, {"history": history, "traj": args.traj, "epochs": history[-1]["epoch"] if history else 0,
                            "seed": args.seed, "eval": res})
    ctx.close()


def eval_main(args):
    xc = yc = args.grid
    ctx = L.Context(air_config(xc, yc))
    ctx.load_state(S.SimState(air_config(xc, yc)))
    sur = Surrogate(xc, yc)
    meta = sur.load(args.weights)
    seeds = range(10 ** 7, 10 ** 7 + args.seeds)
    res = {"trained": evaluate(sur, ctx, seeds, energy_ratio=args.energy_ratio),
           "persistence": evaluate(sur, ctx, seeds, persistence=True),
           "timing_us": timing(sur, ctx)}
    sur.init_params(1)
    res["untrained"] = evaluate(sur, ctx, seeds, check=False)
    res["meta_epochs"] = meta.get("epochs")
    print(json.dumps(res))
    ctx.close()


def calls_main(args):
    """`calls` unchecked surrogate calls and as many 8-tick runs of the plain
    air step on one scene (a workload for rocprofv3)."""
    xc = yc = args.grid
    ctx = L.Context(air_config(xc, yc))
    ctx.load_state(S.SimState(air_config(xc, yc)))
    sur = Surrogate(xc, yc)
    sur.load(args.weights)
    sur.scene(ctx, random_scene(12345, xc, yc))
    for _ in range(args.calls):
        sur.step_air(ctx, check=False)
    ctx.synchronize()
    for _ in range(args.calls):
        sur.plain_ticks(ctx, TICKS_PER_CALL)
    ctx.synchronize()
    print(f"{args.calls} surrogate calls and {args.calls} x {TICKS_PER_CALL} plain air ticks", flush=True)
    ctx.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--grid", type=int, default=512)
    t.add_argument("--traj", type=int, default=2000)
    t.add_argument("--epochs", type=int, default=6)
    t.add_argument("--accum", type=int, default=4)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--warmup-steps", type=int, default=500)
    t.add_argument("--clip", type=float, default=1.0)
    t.add_argument("--seed", type=int, default=7)
    t.add_argument("--eval-seeds", type=int, default=0)
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--init", default=None, help="continue from a saved network (its scales are kept)")
    t.add_argument("--rollout", type=int, default=1,
                   help="train every pair also from the network's own first 1 .. R-1 steps (R-step rollouts)")
    e = sub.add_parser("eval")
    e.add_argument("--grid", type=int, default=512)
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--seeds", type=int, default=256)
    e.add_argument("--energy-ratio", type=float, default=1.05)
    c = sub.add_parser("calls")
    c.add_argument("--grid", type=int, default=512)
    c.add_argument("--weights", default=DEFAULT_WEIGHTS)
    c.add_argument("--calls", type=int, default=30)
    args = ap.parse_args(argv)
    return {"train": train_main, "eval": eval_main, "calls": calls_main}[args.cmd](args)


if __name__ == "__main__":
    sys.exit(main())
