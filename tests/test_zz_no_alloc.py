"""The allocation-free matvec (tnx_heff1_matvec_noalloc) against the
workspace form (tnx_heff1_matvec), through the C-ABI: two calls in a row on
the shared device scratch array, the second with a smaller input, each
bitwise equal to the workspace form (same kernels, tiles and split-K rule);
the workspace form itself against the oracle; sizes above the fixed scratch
rejected with an error."""

import numpy as np
import pytest
import torch

from oracle import dmrg_ref as o

pytestmark = pytest.mark.gpu


def _lib():
    from tnpy_amd import _lib

    return _lib.lib()


def _problem(chil, D, d, chir, seed):
    rng = np.random.default_rng(seed)
    Lh = rng.standard_normal((chil, D, chil))
    Rh = rng.standard_normal((chir, D, chir))
    Wh = o.xxz_bulk(0.7)[:D, :D]
    xh = rng.standard_normal((chil, d, chir))
    return Lh, Wh, Rh, xh


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_noalloc_matches_workspace_form(gpu, precision):
    lib = _lib()
    s = torch.cuda.current_stream().cuda_stream
    print(f"noalloc matvec, precision {precision}", flush=True)
    for chil, chir, seed in ((256, 256, 1), (96, 64, 2)):  # second call smaller
        D, d = 5, 2
        Lh, Wh, Rh, xh = _problem(chil, D, d, chir, seed)
        Lt, Wt, Rt, xt = (torch.as_tensor(a, device="cuda").contiguous() for a in (Lh, Wh, Rh, xh))
        y0 = torch.full_like(xt, np.nan)
        y1 = torch.full_like(xt, np.nan)
        nb = lib.tnx_heff1_workspace_bytes(chil, D, D, d, chir)
        ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
        st = lib.tnx_heff1_matvec(Lt.data_ptr(), Wt.data_ptr(), Rt.data_ptr(), xt.data_ptr(), y0.data_ptr(),
                                  chil, D, D, d, chir, precision, ws.data_ptr(), ws.numel(), s)
        assert st == 0, lib.tnx_last_error()
        st = lib.tnx_heff1_matvec_noalloc(Lt.data_ptr(), Wt.data_ptr(), Rt.data_ptr(), xt.data_ptr(),
                                          y1.data_ptr(), chil, D, D, d, chir, precision, s)
        assert st == 0, lib.tnx_last_error()
        torch.cuda.synchronize()
        a, b = y0.cpu().numpy(), y1.cpu().numpy()
        assert np.array_equal(a, b), f"chil={chil} chir={chir}: max |diff| {np.abs(a - b).max()}"
        ref = o.heff_matvec(Lh, Wh, Rh, xh)
        rel = np.abs(b.reshape(-1) - ref).max() / np.abs(ref).max()
        assert rel < {0: 1e-13, 1: 2e-6, 2: 2e-5}[precision], rel
        del ws


def test_noalloc_rejects_oversize(gpu):
    lib = _lib()
    cap = lib.tnx_noalloc_scratch_bytes()
    assert cap == 64 << 20
    big = 1024
    assert lib.tnx_heff1_workspace_bytes(big, 5, 5, 2, big) > cap
    t = torch.zeros(1, dtype=torch.float64, device="cuda")
    p = t.data_ptr()
    st = lib.tnx_heff1_matvec_noalloc(p, p, p, p, p, big, 5, 5, 2, big, 0, torch.cuda.current_stream().cuda_stream)
    assert st == -1
    assert b"scratch" in lib.tnx_last_error()
