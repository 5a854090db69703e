"""The learned air surrogate's kernels and training (tpt_sur_* in the C ABI).

* Gradient check: the HIP forward and every layer's weight and bias gradient
  against a CPU float32 autograd (torch) of the same U-Net.  Activations are
  rounded to bf16 where the kernels store them (straight through for the
  gradient).  The kernels also keep the intermediate gradients in bf16, so
  the tolerance is a relative L2 of 5e-2 with a cosine above 0.995 per layer.
* Training on plain-solver trajectories lowers the loss.
* The energy check: a network that blows the state up falls back to 8
  plain air ticks, bit for bit.
* The trained network (when its weights are present) beats the untrained
  network and holding the state still, on held-out scenes.
"""
import os

import numpy as np
import pytest

from tpt_hip import lib as L
from tpt_hip import scene as S
from tpt_hip import surrogate as SU

pytestmark = pytest.mark.gpu
SPEC = [(8, 32, 0, 1), (32, 64, 1, 1), (64, 64, 1, 1), (64, 128, 2, 1), (128, 128, 2, 1), (128, 64, 2, 1),
        (64, 64, 1, 1), (64, 32, 1, 1), (32, 32, 0, 1), (32, 8, 0, 0)]


def ceil(v, m):
    return (v + m - 1) // m * m


def layout():
    out, off = [], 0
    for cin, cout, lvl, gelu in SPEC:
        kpad, rows = ceil(9 * cin + 1, 64), ceil(cout, 64)
        out.append({"cin": cin, "cout": cout, "lvl": lvl, "gelu": gelu, "kpad": kpad, "rows": rows, "off": off})
        off += rows * kpad
    return out, off


def bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_round(x):
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32)


def torch_reference(params, x0, H, W, g_out):
    """Forward and weight gradients of the U-Net in torch float32 (CPU)."""
    import torch
    import torch.nn.functional as F

    torch.set_grad_enabled(True)
    lay, _ = layout()

    def rb(t):  # bf16 store, straight-through gradient
        u = t.detach().contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        r = (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16) & 0xFFFFFFFF
        r = torch.where(r >= 2 ** 31, r - 2 ** 32, r).to(torch.int32).view(torch.float32)
        return t + (r - t).detach()

    def gelu(z):
        return 0.5 * z * (1.0 + torch.tanh(0.7978845608 * (z + 0.044715 * z * z * z)))

    mats = []
    for ly in lay:
        m = torch.tensor(params[ly["off"]:ly["off"] + ly["rows"] * ly["kpad"]].reshape(ly["rows"], ly["kpad"]),
                         requires_grad=True)
        mats.append(m)

    def conv(i, x):
        ly = lay[i]
        cin, cout = ly["cin"], ly["cout"]
        Wm = mats[i][:cout]
        w = Wm[:, :9 * cin].reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
        z = F.conv2d(x, w, Wm[:, 9 * cin], padding=1)
        return rb(gelu(z)) if ly["gelu"] else z

    pool = lambda t: rb(F.avg_pool2d(t, 2))  # noqa: E731
    up_add = lambda lo, skip: rb(F.interpolate(lo, scale_factor=2, mode="nearest") + skip)  # noqa: E731
    x = torch.tensor(x0.reshape(H, W, 8)).permute(2, 0, 1)[None]
    a1 = conv(0, x)
    a2 = conv(1, pool(a1))
    a2b = conv(2, a2)
    a3 = conv(3, pool(a2b))
    a3b = conv(4, a3)
    d2a = conv(5, a3b)
    d2b = conv(6, up_add(d2a, a2b))
    d1a = conv(7, d2b)
    d1b = conv(8, up_add(d1a, a1))
    out = conv(9, d1b)
    g = torch.tensor(g_out.reshape(H, W, 8)).permute(2, 0, 1)[None]
    (out * g).sum().backward()
    grads = np.zeros(sum(ly["rows"] * ly["kpad"] for ly in lay), np.float32)
    for ly, m in zip(lay, mats):
        grads[ly["off"]:ly["off"] + m.numel()] = m.grad.numpy().reshape(-1)
    return out.detach()[0].permute(1, 2, 0).numpy().reshape(-1), grads


@pytest.mark.parametrize("H,W", [(32, 64), (40, 48)])
def test_forward_and_weight_gradients_match_torch(gpu, H, W):
    rng = np.random.default_rng(H * 1000 + W)
    lay, n = layout()
    sur = SU.Surrogate(W, H)
    assert sur.n_params == n
    params = np.zeros(n, np.float32)
    for ly in lay:
        m = params[ly["off"]:ly["off"] + ly["rows"] * ly["kpad"]].reshape(ly["rows"], ly["kpad"])
        sd = np.sqrt(2.0 / (9 * ly["cin"]))
        m[:ly["cout"], :9 * ly["cin"] + 1] = rng.normal(0.0, sd, (ly["cout"], 9 * ly["cin"] + 1))
    params = bf16_round(params)
    sur.set_params(params)
    x0 = bf16_round(rng.normal(0.0, 1.0, H * W * 8).astype(np.float32))
    g_out = np.zeros((H * W, 8), np.float32)
    g_out[:, :4] = rng.normal(0.0, 1.0, (H * W, 4))
    g_out = bf16_round(g_out.reshape(-1))
    print(f"[surrogate grad check] {H}x{W}", flush=True)
    out = sur.forward_raw(bf16_bits(x0))
    sur.backward_raw(bf16_bits(g_out))
    grads = sur.grads()
    ref_out, ref_grads = torch_reference(params, x0, H, W, g_out)
    err = np.abs(out - ref_out).max() / np.abs(ref_out).max()
    assert err < 2e-2, err
    for i, ly in enumerate(lay):
        sl = slice(ly["off"], ly["off"] + ly["rows"] * ly["kpad"])
        a, b = grads[sl].astype(np.float64), ref_grads[sl].astype(np.float64)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        bias_a = a.reshape(ly["rows"], ly["kpad"])[:ly["cout"], 9 * ly["cin"]]
        bias_b = b.reshape(ly["rows"], ly["kpad"])[:ly["cout"], 9 * ly["cin"]]
        print(f"  layer {i}: rel {rel:.3e} cos {cos:.6f}")
        assert rel < 5e-2 and cos > 0.995, (i, rel, cos)
        assert np.linalg.norm(bias_a - bias_b) <= 5e-2 * np.linalg.norm(bias_b) + 1e-12, i
        pad = a.reshape(ly["rows"], ly["kpad"]).copy()
        pad[:ly["cout"], :9 * ly["cin"] + 1] = 0.0
        assert not pad.any(), i   # padding rows and columns get no gradient
    sur.close()


def small_world(xc=64, yc=64):
    cfg = SU.air_config(xc, yc)
    ctx = L.Context(cfg)
    ctx.load_state(S.SimState(cfg))
    return ctx, SU.Surrogate(xc, yc)


@pytest.mark.parametrize("xc,yc", [(64, 64), (96, 40)])
def test_fused_inference_matches_separate_kernels(gpu, xc, yc):
    """The inference path with the pools, upsample-adds and input
    preparation folded into the convolutions gives the bits of the
    separate kernels, over three calls."""
    ctx, sur = small_world(xc, yc)
    sur.init_params(5)
    desc = SU.random_scene(77, xc, yc)
    n = xc * yc
    got = []
    for fused in (False, True):
        sur.set_fused(fused)
        sur.scene(ctx, desc)
        for _ in range(3):
            sur.step_air(ctx, check=False)
        got.append(SU.air_state(ctx, n))
    for k in got[0]:
        assert np.array_equal(got[0][k].view(np.uint32), got[1][k].view(np.uint32)), k
    assert np.isfinite(got[1]["pv"]).all()
    ctx.close()
    sur.close()


def test_training_lowers_the_loss(gpu):
    ctx, sur = small_world()
    sur.set_scales(SU.IN_SCALES, SU.delta_scales(ctx, sur, range(900, 903)))
    SU.build_dataset(sur, ctx, list(range(16)))
    pairs = np.array([(t, k) for t in range(16) for k in range(8)], np.uint32)
    before = sur.eval_loss(pairs)
    print(f"[surrogate training] loss before {before:.4e}", flush=True)
    rng = np.random.default_rng(0)
    for ep in range(12):
        sur.train(pairs[rng.permutation(len(pairs))], 4, 2e-3)
    after = sur.eval_loss(pairs)
    print(f"  after {after:.4e}", flush=True)
    assert np.isfinite(after) and after < 0.5 * before
    ctx.close()
    sur.close()


def test_two_step_rollout_training(gpu):
    """tpt_sur_set_rollout(2): pairs train from the network's own first
    step; a pair without a snapshot two calls ahead is refused, a zero
    learning rate leaves the parameters alone, and training lowers the
    two-step loss.  The same for three steps; nine are refused."""
    ctx, sur = small_world()
    sur.set_scales(SU.IN_SCALES, SU.delta_scales(ctx, sur, range(900, 903)))
    SU.build_dataset(sur, ctx, list(range(16)))
    sur.set_rollout(2)
    with pytest.raises(RuntimeError):
        sur.train(np.array([(0, 7)], np.uint32), 1, 0.0)      # snapshot 9 does not exist
    pairs = np.array([(t, k) for t in range(16) for k in range(7)], np.uint32)
    p0 = sur.params()
    before = sur.train(pairs, 4, 0.0)
    assert np.array_equal(sur.params(), p0)
    print(f"[surrogate rollout] two-step loss before {before:.4e}", flush=True)
    rng = np.random.default_rng(1)
    for ep in range(12):
        sur.train(pairs[rng.permutation(len(pairs))], 4, 2e-3)
    after = sur.train(pairs, 4, 0.0)
    print(f"  after {after:.4e}", flush=True)
    assert np.isfinite(after) and after < 0.5 * before
    # deeper rollouts: 3 steps need snapshot k+3, the horizon stops at 8
    sur.set_rollout(3)
    with pytest.raises(RuntimeError):
        sur.train(np.array([(0, 6)], np.uint32), 1, 0.0)
    with pytest.raises(RuntimeError):
        sur.set_rollout(9)
    three = np.array([(t, k) for t in range(16) for k in range(6)], np.uint32)
    p1 = sur.params()
    before3 = sur.train(three, 4, 0.0)
    assert np.array_equal(sur.params(), p1)
    for ep in range(6):
        sur.train(three[rng.permutation(len(three))], 4, 1e-3)
    after3 = sur.train(three, 4, 0.0)
    print(f"  three-step loss {before3:.4e} -> {after3:.4e}", flush=True)
    assert np.isfinite(after3) and after3 < before3
    sur.set_rollout(1)
    ctx.close()
    sur.close()


def test_energy_check_falls_back_to_plain_ticks_bit_exact(gpu):
    ctx, sur = small_world()
    p = sur.params() * 0.0
    lay, _ = layout()
    head = lay[-1]
    m = p[head["off"]:head["off"] + head["rows"] * head["kpad"]].reshape(head["rows"], head["kpad"])
    m[:4, 9 * head["cin"]] = 50.0          # a constant change of 50 scale units per call: energy explodes
    sur.set_params(p)
    desc = SU.random_scene(4242, 64, 64)
    n = 64 * 64
    sur.scene(ctx, desc)
    sur.plain_ticks(ctx, 8)
    want = SU.air_state(ctx, n)
    sur.scene(ctx, desc)
    assert sur.step_air(ctx, check=True) is True
    got = SU.air_state(ctx, n)
    for k in want:
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k
    sur.scene(ctx, desc)
    assert sur.step_air(ctx, check=False) is False     # unchecked: the prediction stays
    got = SU.air_state(ctx, n)
    assert not np.array_equal(got["pv"], want["pv"])
    ctx.close()
    sur.close()


@pytest.mark.skipif(not os.path.exists(SU.DEFAULT_WEIGHTS), reason="no trained weights in the package")
def test_trained_network_beats_untrained_and_persistence(gpu):
    cfg = SU.air_config(512, 512)
    ctx = L.Context(cfg)
    ctx.load_state(S.SimState(cfg))
    sur = SU.Surrogate(512, 512)
    sur.load(SU.DEFAULT_WEIGHTS)
    seeds = range(10 ** 7, 10 ** 7 + 8)
    trained = SU.evaluate(sur, ctx, seeds)
    still = SU.evaluate(sur, ctx, seeds, persistence=True)
    print(f"[surrogate eval] trained {trained} persistence {still['rel_l2']}", flush=True)
    assert trained["rel_l2"]["mean"] < 0.5 * still["rel_l2"]["mean"]
    sur.init_params(3)
    untrained = SU.evaluate(sur, ctx, seeds, check=False)
    assert trained["rel_l2"]["mean"] < 0.5 * untrained["rel_l2"]["mean"]
    ctx.close()
    sur.close()
