"""The allocation-free gravity solve (tpt_gravity_solve_noalloc) against the
context's solve (tpt_gravity_solve): two calls in a row on one GPU, the
second on a smaller grid.  Each must equal the context's forces bit for bit
(the same kernels on the same padded grid) and stay within 1e-4 of the FP64
solve; oversized grids are refused before any GPU work."""
import numpy as np
import pytest

from oracle import oracle as O
from tpt_hip import lib as L
from tpt_hip import scene as S

pytestmark = pytest.mark.gpu
TOL = 1e-4


def mass_state(xres, yres, seed, grav_box):
    st = S.SimState(S.default_config(xres, yres, capacity=65536))
    rng = np.random.default_rng(seed)
    n = st.xcells * st.ycells
    st.arrays["grav_mass_prev"][:] = ((rng.uniform(-20, 100, n) * 65536).astype(np.int32) * (rng.random(n) < 0.2))
    wall = st.arrays["wall_bmap"].reshape(st.ycells, st.xcells)
    if grav_box:
        y0, x0 = st.ycells // 3, st.xcells // 3
        wall[y0, x0:x0 + 20] = wall[y0 + 15, x0:x0 + 20] = 14
        wall[y0:y0 + 16, x0] = wall[y0:y0 + 16, x0 + 19] = 14
    st.arrays["grav_mask"][:] = O.gravity_mask(st.arrays["wall_bmap"], st.xcells, st.ycells)
    return st


def context_solve(st):
    n = st.xcells * st.ycells
    with L.Context(st.cfg) as ctx:
        ctx.load_state(st)
        ctx.gravity_solve()
        return ctx.download("grav_fx", np.float32, n), ctx.download("grav_fy", np.float32, n)


def noalloc_solve(st):
    import torch  # device buffers for the caller's side of the call only
    dev = torch.device("cuda:0")
    n = st.xcells * st.ycells
    mass = torch.from_numpy(st.arrays["grav_mass_prev"].copy()).to(dev)
    mask = torch.from_numpy(st.arrays["grav_mask"].view(np.int32).copy()).to(dev)
    fx = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    fy = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    rc = L.load().tpt_gravity_solve_noalloc(mass.data_ptr(), mask.data_ptr(), fx.data_ptr(), fy.data_ptr(),
                                            st.xcells, st.ycells, stream)
    assert rc == 0, L.load().tpt_last_error()
    torch.cuda.synchronize()
    return fx.cpu().numpy(), fy.cpu().numpy()


def rel_err(got, want):
    scale = np.abs(want).max()
    return float(np.abs(got.astype(np.float64) - want).max() / scale)


def test_noalloc_solve_matches_context_solve_twice(gpu):
    # the largest grid it takes (512x512 cells, FFT 1024x1024), then a smaller one
    for (X, Y, seed, box) in ((2048, 2048, 3, True), (612, 384, 4, False)):
        print(f"[noalloc] {X}x{Y}", flush=True)
        st = mass_state(X, Y, seed, box)
        fx0, fy0 = context_solve(st)
        fx1, fy1 = noalloc_solve(st)
        assert np.array_equal(fx1.view(np.uint32), fx0.view(np.uint32))
        assert np.array_equal(fy1.view(np.uint32), fy0.view(np.uint32))
        fx64, fy64 = O.gravity_solve_f64(st)
        assert rel_err(fx1, fx64) < TOL and rel_err(fy1, fy64) < TOL


def test_noalloc_solve_refuses_oversized_grids(gpu):
    lib = L.load()
    # 600 cells need a 2048-wide FFT grid: refused on the host, before any launch
    rc = lib.tpt_gravity_solve_noalloc(8, 8, 8, 8, 600, 100, None)
    assert rc == L.TPT_EINVAL and b"at most" in lib.tpt_last_error()
    assert lib.tpt_gravity_solve_noalloc(None, 8, 8, 8, 10, 10, None) == L.TPT_EINVAL
    assert lib.tpt_gravity_solve_noalloc(8, 8, 8, 8, 0, 10, None) == L.TPT_EINVAL
