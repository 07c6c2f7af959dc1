"""ShiftInvertDMRG on the GPU (tnpy/finite_dmrg.py:266-407), pinned as the
reference pins it (tests/test_finite_dmrg.py:26-72: RandomHeisenberg n=10,
h=10.5, seed=2022, bond_dim=64, offsets -0.1, 0.1, 0.2; energy, restored
eigenvector up to sign and restored energy within 1e-6 of exact
diagonalisation), and its two generalized eigensolvers against
numpy/scipy-free dense references."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import dmrg_ref as o

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def quiet_logger():
    import logging

    lg = logging.getLogger("tnpy")
    old = lg.level
    lg.setLevel(logging.WARNING)
    yield
    lg.setLevel(old)


def _dense_generalized_lowest(A, B):
    """Lowest pair of A x = l B x by the Cholesky reduction (reference)."""
    Lc = np.linalg.cholesky(B)
    Li = np.linalg.inv(Lc)
    w, V = np.linalg.eigh(Li @ A @ Li.T)
    return w[0], Li.T @ V[:, 0]


@pytest.mark.parametrize("n", [1, 7, 64, 199])
def test_sygv_lowest(gpu, n):
    from tnpy_amd import _lib

    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    B = (Q * np.linspace(0.05, 3.0, n)) @ Q.T
    lib = _lib.lib()
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    ws = torch.empty(lib.tnx_sygv_lowest_workspace_bytes(n), dtype=torch.uint8, device="cuda")
    ev = C.c_double(0)
    Ad = torch.as_tensor(A, device="cuda")
    Bd = torch.as_tensor(B, device="cuda")
    st = lib.tnx_sygv_lowest(Ad.data_ptr(), Bd.data_ptr(), n, x.data_ptr(), C.byref(ev), ws.data_ptr(), ws.numel(),
                             torch.cuda.current_stream().cuda_stream)
    assert st == 0, lib.tnx_last_error()
    xh = x.cpu().numpy()
    lam, ref = _dense_generalized_lowest(A, B)
    assert abs(ev.value - lam) <= 1e-11 * max(1.0, abs(lam))
    assert abs(xh @ B @ xh - 1.0) < 1e-11
    assert np.linalg.norm(A @ xh - ev.value * B @ xh) <= 1e-10 * np.abs(A).max() * np.sqrt(n)


def test_generalized_davidson_matches_dense(gpu):
    """The Krylov solver on a one-site problem of (H - eps), M = (H - eps)^2,
    against the dense generalized solution of the same local matrices."""
    from tnpy_amd import Environment, MatrixProductState, _lib
    from tnpy_amd.model import RandomHeisenberg

    n, chi, site = 10, 16, 4
    model = RandomHeisenberg(n=n, h=10.5, seed=2022, offset=0.1)
    mps = MatrixProductState(o.random_mps(n, chi, 2, seed=3))
    from tnpy_amd import Direction

    for s in range(site):
        mps.split_tensor(s, Direction.RIGHTWARD)
    env = Environment(model.mpo, mps, build_left=True)
    env2 = Environment(model.mpo.square(), mps, build_left=True)
    A = env.one_site_full_matrix(site)
    B = env2.one_site_full_matrix(site)
    lam, _ = _dense_generalized_lowest(A, B)
    l, Dl, Dr, d, r = env.site_dims(site)
    _, D2l, D2r, _, _ = env2.site_dims(site)
    N = l * d * r
    lib = _lib.lib()
    opts = _lib.EigshOpts(tol=1e-12, max_basis=32, max_matvecs=4000, precision=0, keep=8)
    info = _lib.EigshInfo()
    vec = torch.empty(N, dtype=torch.float64, device="cuda")
    hvec = torch.empty(N, dtype=torch.float64, device="cuda")
    ws = torch.empty(lib.tnx_eigsh_gen_heff1_workspace_bytes(l, Dl, Dr, D2l, D2r, d, r, 32), dtype=torch.uint8,
                     device="cuda")
    v0 = mps.tensors[site].reshape(-1).contiguous()
    st = lib.tnx_eigsh_gen_heff1(env.left[site].data_ptr(), env._W[site].data_ptr(), env.right[site].data_ptr(),
                                 env2.left[site].data_ptr(), env2._W[site].data_ptr(), env2.right[site].data_ptr(),
                                 v0.data_ptr(), vec.data_ptr(), hvec.data_ptr(), l, Dl, Dr, D2l, D2r, d, r,
                                 C.byref(opts), C.byref(info), ws.data_ptr(), ws.numel(),
                                 torch.cuda.current_stream().cuda_stream)
    assert st == 0, lib.tnx_last_error()
    x = vec.cpu().numpy()
    print(f"N={N}: lambda {info.eval:.12f} vs dense {lam:.12f}, matvecs {info.matvecs}", flush=True)
    assert abs(info.eval - lam) <= 1e-9 * abs(lam)
    assert abs(x @ B @ x - 1.0) < 1e-10
    np.testing.assert_allclose(hvec.cpu().numpy(), A @ x, atol=1e-10 * np.abs(A @ x).max())


@pytest.mark.parametrize("offset", [-0.1, 0.1, 0.2])
def test_shift_invert_vs_ed(gpu, offset):
    # tests/test_finite_dmrg.py:26-72 of the reference
    from tnpy_amd import ExactDiagonalization, ShiftInvertDMRG
    from tnpy_amd.model import RandomHeisenberg

    model = RandomHeisenberg(n=10, h=10.5, seed=2022)
    ed = ExactDiagonalization(model.mpo)
    idx = ed.evals[np.where(ed.evals < offset)].argmax()
    nearest_eval, nearest_evec = ed.evals[idx], ed.evecs[:, idx]
    shifted = RandomHeisenberg(n=model.n, h=model.h, seed=model.seed, offset=offset)
    sidmrg = ShiftInvertDMRG(shifted.mpo, bond_dim=2**6, offset=offset, seed=1)
    energies = sidmrg.run(tol=1e-8)
    print(f"offset {offset}: E={energies[-1]:.12f} vs ED {nearest_eval:.12f}, sweeps {len(energies)}", flush=True)
    np.testing.assert_allclose(energies[-1], nearest_eval, atol=1e-6)
    restored = sidmrg.restored_mps.to_dense()
    if not np.allclose(restored, nearest_evec, atol=1e-6):
        np.testing.assert_allclose(-1 * restored, nearest_evec, atol=1e-6)
    np.testing.assert_allclose(energies[-1], sidmrg.measurements.expectation_value(model.mpo), atol=1e-6)
