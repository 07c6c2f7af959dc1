"""GPU kernels of libtnx against the CPU oracle, through the C-ABI.

Tolerances: FP64 paths to ~1e-13 relative; FP32 / BF16x3 contractions
within their arithmetic's rounding (stated per test); iterative solvers to
their residual criterion.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import dmrg_ref as o

pytestmark = pytest.mark.gpu


_KEEP = []


def dev(a):
    """Device copy that stays alive until the test ends (a bare
    ``dev(x).data_ptr()`` would let the caching allocator reuse the block
    before the kernel reads it)."""
    t = torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    _KEEP.append(t)
    return t


@pytest.fixture(autouse=True)
def _release():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def L():
    from tnpy_amd import _lib

    return _lib.lib()


def ws(nbytes):
    buf = torch.empty(int(nbytes) + 256, dtype=torch.uint8, device="cuda")
    _KEEP.append(buf)
    return buf, buf.data_ptr(), buf.numel()


def stream():
    return torch.cuda.current_stream().cuda_stream


def check(st, what="call"):
    assert st == 0, f"{what}: status {st}: {L().tnx_last_error()}"


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}
# relative Frobenius error bounds of one contraction in each arithmetic
TOL = {"fp64": 1e-13, "fp32": 3e-6, "bf16x3": 1e-4}


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("ak, bn", [(1, 1), (1, 0), (0, 1), (0, 0)])
@pytest.mark.parametrize("M, N, K", [(1, 1, 1), (17, 33, 9), (130, 70, 260), (256, 384, 512)])
def test_gemm(gpu, prec, ak, bn, M, N, K):
    print(f"gemm {prec} ak={ak} bn={bn} {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    Cin = rng.standard_normal((M, N))
    Ad = dev(A if ak else A.T)
    Bd = dev(B if bn else B.T)
    Cd = dev(Cin)
    lda = K if ak else M
    ldb = N if bn else K
    check(L().tnx_gemm_f64(M, N, K, 1, 0.75, Ad.data_ptr(), lda, 0, ak, Bd.data_ptr(), ldb, 0, bn,
                           -0.5, Cd.data_ptr(), N, 0, PREC[prec], stream()), "gemm")
    ref = 0.75 * A @ B - 0.5 * Cin
    assert relerr(host(Cd), ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("ak, bn", [(1, 1), (1, 0), (0, 1), (0, 0)])
def test_gemm_typed_matches_generic(gpu, prec, ak, bn):
    """An operand at a 16-byte aligned address takes the typed staging
    (16-byte buffer loads); the same operand one element further takes the
    generic staging. Both must give the same bits (ragged edge tiles and a
    partial last k-tile included)."""
    print(f"gemm typed/generic {prec} ak={ak} bn={bn}", flush=True)
    M, N, K = 130, 70, 260
    rng = np.random.default_rng(11)
    A = rng.standard_normal((M, K))
    B = rng.standard_normal((K, N))
    Ah = np.ascontiguousarray(A if ak else A.T)
    Bd = dev(np.ascontiguousarray(B if bn else B.T))
    lda = K if ak else M
    ldb = N if bn else K

    def run(off):
        buf = torch.zeros(Ah.size + 2, dtype=torch.float64, device="cuda")
        buf[off:off + Ah.size] = torch.from_numpy(Ah.ravel()).cuda()
        Cd = torch.zeros(M, N, dtype=torch.float64, device="cuda")
        check(L().tnx_gemm_f64(M, N, K, 1, 1.0, buf.data_ptr() + 8 * off, lda, 0, ak, Bd.data_ptr(), ldb, 0,
                               bn, 0.0, Cd.data_ptr(), N, 0, PREC[prec], stream()), "gemm")
        return host(Cd)

    typed, generic = run(0), run(1)
    assert relerr(typed, A @ B) < TOL[prec]
    assert np.array_equal(typed, generic)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("bt, M, N, K", [(1, 1030, 1950, 40), (1, 2000, 1000, 24), (3, 512, 512, 36),
                                         (3, 520, 1000, 20), (1, 512, 512, 4096)])
def test_gemm_large_grids(gpu, prec, bt, M, N, K):
    """Grids of 64 or more output tiles run in the XCD-aware tile order:
    tile counts that are and are not multiples of 8, batches and a K split.
    C starts as non-zero data with beta = -0.5, so a tile left unwritten or
    written from the wrong place fails the comparison."""
    print(f"gemm large grid {prec} bt={bt} {M}x{N}x{K}", flush=True)
    rng = np.random.default_rng(bt * 1000 + M + N + K)
    A = rng.standard_normal((bt, M, K))
    B = rng.standard_normal((bt, K, N))
    Cin = rng.standard_normal((bt, M, N))
    Cd = dev(Cin)
    check(L().tnx_gemm_f64(M, N, K, bt, 1.0, dev(A).data_ptr(), K, M * K, 1, dev(B).data_ptr(), N, K * N, 1,
                           -0.5, Cd.data_ptr(), N, M * N, PREC[prec], stream()), "gemm")
    assert relerr(host(Cd), A @ B - 0.5 * Cin) < TOL[prec]


def test_gemm_batched(gpu):
    rng = np.random.default_rng(5)
    bt, M, N, K = 5, 40, 24, 33
    A = rng.standard_normal((bt, M, K))
    B = rng.standard_normal((bt, K, N))
    out = torch.zeros(bt, M, N, dtype=torch.float64, device="cuda")
    check(L().tnx_gemm_f64(M, N, K, bt, 1.0, dev(A).data_ptr(), K, M * K, 1, dev(B).data_ptr(), N, K * N, 1,
                           0.0, out.data_ptr(), N, M * N, 0, stream()))
    assert relerr(host(out), A @ B) < 1e-13


# ---------------------------------------------------------------------------
def rand_env(rng, chil, Dl, Dr, d, chir):
    Lt = rng.standard_normal((chil, Dl, chil))
    Rt = rng.standard_normal((chir, Dr, chir))
    W = rng.standard_normal((Dl, Dr, d, d))
    x = rng.standard_normal((chil, d, chir))
    return Lt, W, Rt, x


SHAPES = [(1, 1, 5, 2, 2), (2, 5, 5, 2, 4), (8, 5, 5, 2, 16), (33, 5, 5, 2, 17), (64, 5, 5, 2, 64),
          (16, 5, 1, 2, 1), (12, 25, 25, 2, 12), (10, 6, 6, 2, 20), (9, 4, 3, 3, 7)]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("shape", SHAPES)
def test_heff1_matvec(gpu, prec, shape):
    chil, Dl, Dr, d, chir = shape
    print(f"matvec {prec} {shape}", flush=True)
    rng = np.random.default_rng(sum(shape))
    Lt, W, Rt, x = rand_env(rng, *shape)
    y = torch.empty(chil * d * chir, dtype=torch.float64, device="cuda")
    _, wp, wb = ws(L().tnx_heff1_workspace_bytes(*shape))
    check(L().tnx_heff1_matvec(dev(Lt).data_ptr(), dev(W).data_ptr(), dev(Rt).data_ptr(), dev(x).data_ptr(),
                               y.data_ptr(), *shape, PREC[prec], wp, wb, stream()), "matvec")
    ref = o.heff_matvec(Lt, W, Rt, x)
    assert relerr(host(y), ref) < 10 * TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_heff1_matvec_split_k(gpu, prec):
    """chi = 256: the stage-3 contraction (K = chi * D = 1280 over a 256 x 512
    output) is cut into four K splits with FP64 partial sums, and every
    operand takes the typed (16-byte buffer load) staging."""
    shape = (256, 5, 5, 2, 256)
    print(f"matvec split-K {prec} {shape}", flush=True)
    rng = np.random.default_rng(256)
    Lt, W, Rt, x = rand_env(rng, *shape)
    y = torch.empty(256 * 2 * 256, dtype=torch.float64, device="cuda")
    _, wp, wb = ws(L().tnx_heff1_workspace_bytes(*shape))
    check(L().tnx_heff1_matvec(dev(Lt).data_ptr(), dev(W).data_ptr(), dev(Rt).data_ptr(), dev(x).data_ptr(),
                               y.data_ptr(), *shape, PREC[prec], wp, wb, stream()), "matvec")
    ref = o.heff_matvec(Lt, W, Rt, x)
    assert relerr(host(y), ref) < 10 * TOL[prec]


@pytest.mark.parametrize("shape", SHAPES)
def test_heff1_dense(gpu, shape):
    chil, Dl, Dr, d, chir = shape
    rng = np.random.default_rng(3 + sum(shape))
    Lt, W, Rt, _ = rand_env(rng, *shape)
    N = chil * d * chir
    M = torch.empty(N, N, dtype=torch.float64, device="cuda")
    check(L().tnx_heff1_dense(dev(Lt).data_ptr(), dev(W).data_ptr(), dev(Rt).data_ptr(), M.data_ptr(), *shape,
                              stream()))
    assert relerr(host(M), o.heff_dense(Lt, W, Rt)) < 1e-14


@pytest.mark.parametrize("side", ["left", "right"])
@pytest.mark.parametrize("shape", SHAPES)
def test_env_update(gpu, side, shape):
    chil, Dl, Dr, d, chir = shape
    print(f"env {side} {shape}", flush=True)
    rng = np.random.default_rng(11 + sum(shape))
    Lt, W, Rt, A = rand_env(rng, *shape)
    _, wp, wb = ws(L().tnx_env_workspace_bytes(*shape))
    if side == "left":
        out = torch.empty(chir, Dr, chir, dtype=torch.float64, device="cuda")
        check(L().tnx_env_left(dev(Lt).data_ptr(), dev(A).data_ptr(), dev(W).data_ptr(), out.data_ptr(), *shape,
                               0, wp, wb, stream()))
        ref = o.env_left(Lt, A, W)
    else:
        out = torch.empty(chil, Dl, chil, dtype=torch.float64, device="cuda")
        check(L().tnx_env_right(dev(Rt).data_ptr(), dev(A).data_ptr(), dev(W).data_ptr(), out.data_ptr(), *shape,
                                0, wp, wb, stream()))
        ref = o.env_right(Rt, A, W)
    assert relerr(host(out), ref) < 1e-13


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4, 16, 64, 100, 199, 256])
def test_syev_lowest(gpu, n):
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    M = X + X.T
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    ev = C.c_double()
    _, wp, wb = ws(L().tnx_syev_lowest_workspace_bytes(n))
    check(L().tnx_syev_lowest(dev(M).data_ptr(), n, v.data_ptr(), C.byref(ev), wp, wb, stream()))
    w = np.linalg.eigvalsh(M)
    nrm = np.abs(w).max()
    assert abs(ev.value - w[0]) <= 1e-13 * nrm
    vv = host(v)
    assert abs(np.linalg.norm(vv) - 1) < 1e-12
    assert np.linalg.norm(M @ vv - ev.value * vv) <= 1e-12 * nrm


def test_syev_degenerate(gpu):
    # ferromagnet-like exact degeneracy of the lowest level
    n = 40
    Q, _ = np.linalg.qr(np.random.default_rng(2).standard_normal((n, n)))
    w = np.concatenate([[-3.0, -3.0, -3.0], np.linspace(-1, 2, n - 3)])
    M = (Q * w) @ Q.T
    v = torch.empty(n, dtype=torch.float64, device="cuda")
    ev = C.c_double()
    _, wp, wb = ws(L().tnx_syev_lowest_workspace_bytes(n))
    check(L().tnx_syev_lowest(dev(M).data_ptr(), n, v.data_ptr(), C.byref(ev), wp, wb, stream()))
    vv = host(v)
    assert abs(ev.value + 3.0) < 1e-13
    assert np.linalg.norm(M @ vv + 3.0 * vv) < 1e-12


# ---------------------------------------------------------------------------
def physical_env(n, chi, site, seed, model="xxz"):
    """Environments of a real DMRG problem (oracle), for solver tests."""
    mpo = o.xxz_mpo(n, 0.5) if model == "xxz" else o.random_heisenberg_mpo(n, 0.5, seed=seed)
    mps = o.random_mps(n, chi, 2, seed=seed)
    # left-canonical to the left of `site`
    for i in range(site):
        o.split_tensor_svd(mps, i, o.RIGHTWARD)
    d = o.OracleDMRG(mpo, mps)
    return d.left[site], mpo[site], d.right[site], mps[site]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("n, chi, site", [(10, 32, 4), (12, 16, 6), (14, 48, 7)])
def test_eigsh(gpu, prec, n, chi, site):
    from tnpy_amd import _lib

    print(f"eigsh {prec} n={n} chi={chi} site={site}", flush=True)
    Lt, W, Rt, A = physical_env(n, chi, site, seed=n + chi)
    M = o.heff_dense(Lt, W, Rt)
    w, V = np.linalg.eigh(M)
    shape = (Lt.shape[0], W.shape[0], W.shape[1], W.shape[2], Rt.shape[0])
    N = A.size
    vec = torch.empty(N, dtype=torch.float64, device="cuda")
    hvec = torch.empty(N, dtype=torch.float64, device="cuda")
    opts = _lib.EigshOpts(tol=1e-10, max_basis=24, max_matvecs=3000, precision=PREC[prec])
    info = _lib.EigshInfo()
    _, wp, wb = ws(L().tnx_eigsh_heff1_workspace_bytes(*shape, 24))
    check(L().tnx_eigsh_heff1(dev(Lt).data_ptr(), dev(W).data_ptr(), dev(Rt).data_ptr(), dev(A).data_ptr(),
                              vec.data_ptr(), hvec.data_ptr(), *shape, C.byref(opts), C.byref(info), wp, wb,
                              stream()), "eigsh")
    nrm = np.abs(w).max()
    assert info.converged == 1
    assert abs(info.eval - w[0]) <= 1e-12 * nrm, (info.eval, w[0], info.matvecs)
    v = host(vec)
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert np.linalg.norm(M @ v - w[0] * v) <= 2e-10 * nrm
    np.testing.assert_allclose(host(hvec), M @ v, atol=1e-11 * nrm)
    if prec != "fp64":
        assert info.matvecs_f64 >= 1


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("m, n", [(1, 1), (7, 3), (37, 23), (64, 64), (512, 256), (600, 300),
                                  (1100, 513), (2048, 1024), (3000, 700), (5000, 300)])
def test_qr(gpu, m, n):
    print(f"qr {m}x{n}", flush=True)
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n))
    Ad = dev(A)
    R = torch.empty(n, n, dtype=torch.float64, device="cuda")
    _, wp, wb = ws(L().tnx_qr_workspace_bytes(m, n))
    check(L().tnx_qr(Ad.data_ptr(), m, n, R.data_ptr(), wp, wb, stream()), "qr")
    Q = host(Ad)
    Rh = host(R)
    assert np.abs(np.tril(Rh, -1)).max() == 0.0
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-13 * np.sqrt(n) * 10
    assert relerr(Q @ Rh, A) < 1e-13 * 10


@pytest.mark.parametrize("m, n, rank, zero_cols", [(2000, 600, 3, True), (2048, 1024, 202, False),
                                                   (2048, 1024, 0, False), (1024, 1024, 40, True)])
def test_qr_rank_deficient(gpu, m, n, rank, zero_cols):
    # rank-deficient inputs (the converged ferromagnetic chain's site tensors
    # have rank ~ n/2 + 1 at chi=1024): Q must stay orthonormal
    print(f"qr rank-deficient {m}x{n} rank {rank}", flush=True)
    rng = np.random.default_rng(1)
    A = rng.standard_normal((m, rank)) @ rng.standard_normal((rank, n)) if rank else np.zeros((m, n))
    if zero_cols:
        A[:, 100:110] = 0.0
    Ad = dev(A)
    R = torch.empty(n, n, dtype=torch.float64, device="cuda")
    _, wp, wb = ws(L().tnx_qr_workspace_bytes(m, n))
    check(L().tnx_qr(Ad.data_ptr(), m, n, R.data_ptr(), wp, wb, stream()), "qr")
    Q = host(Ad)
    Rh = host(R)
    assert np.abs(np.tril(Rh, -1)).max() == 0.0
    assert np.linalg.norm(Q.T @ Q - np.eye(n)) < 1e-11
    assert np.abs(Q @ Rh - A).max() <= 1e-12 * max(1.0, np.abs(A).max())


def graded(rng, m, n, span):
    """m x n with singular values 10^0 .. 10^-span (log-spaced) and random
    singular vectors, the spectrum shape of converged DMRG site tensors."""
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    sv = np.logspace(0, -span, n)
    return (U * sv) @ V.T, sv


@pytest.mark.parametrize("m, n, span", [(2048, 1024, 14), (2048, 1024, 8), (1536, 768, 16), (2048, 1000, 12)])
def test_qr_graded(gpu, m, n, span):
    """Panels of a graded matrix carry condition numbers up to 10^span:
    Q stays orthonormal, Q R reproduces A to rounding, and R keeps A's
    singular values to ~eps ||A|| (the grouped Q formation runs for
    n = 1024 and 768, the panel-by-panel one for n = 1000)."""
    print(f"qr graded {m}x{n} span {span}", flush=True)
    rng = np.random.default_rng(m + n + span)
    A, sv = graded(rng, m, n, span)
    Ad = dev(A)
    R = torch.empty(n, n, dtype=torch.float64, device="cuda")
    _, wp, wb = ws(L().tnx_qr_workspace_bytes(m, n))
    check(L().tnx_qr(Ad.data_ptr(), m, n, R.data_ptr(), wp, wb, stream()), "qr")
    Q = host(Ad)
    Rh = host(R)
    assert np.abs(np.tril(Rh, -1)).max() == 0.0
    assert np.all(np.diag(Rh) >= 0.0)
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
    assert relerr(Q @ Rh, A) < 1e-14
    assert np.abs(np.linalg.svd(Rh, compute_uv=False) - sv).max() < 1e-13


@pytest.mark.parametrize("right", [True, False])
def test_gauge_moves_chi1024(gpu, right):
    """Both gauge moves on a 1024 x 2 x 1024 site tensor with a graded
    spectrum: the moved site is an isometry and the two-site product is
    unchanged."""
    print(f"gauge {'right' if right else 'left'} chi=1024", flush=True)
    chi, d = 1024, 2
    rng = np.random.default_rng(7 if right else 8)
    if right:
        A = graded(rng, chi * d, chi, 12)[0].reshape(chi, d, chi)
        nxt = rng.standard_normal((chi, d, 4))
        gp, wp, wb = ws(L().tnx_gauge_workspace_bytes(chi, d, chi))
        Ad, out = dev(A), torch.empty(chi, d, 4, dtype=torch.float64, device="cuda")
        check(L().tnx_gauge_right(Ad.data_ptr(), dev(nxt).data_ptr(), out.data_ptr(), chi, d, chi, d * 4,
                                  wp, wb, stream()), "gauge_right")
        Q = host(Ad).reshape(chi * d, chi)
        assert np.abs(Q.T @ Q - np.eye(chi)).max() < 1e-13
        before = A.reshape(chi * d, chi) @ nxt.reshape(chi, -1)
        after = Q @ host(out).reshape(chi, -1)
    else:
        A = graded(rng, chi * d, chi, 12)[0].T.reshape(chi, d, chi)
        prv = rng.standard_normal((4, d, chi))
        gp, wp, wb = ws(L().tnx_gauge_workspace_bytes(chi, d, chi))
        Ad, out = dev(A), torch.empty(4, d, chi, dtype=torch.float64, device="cuda")
        check(L().tnx_gauge_left(Ad.data_ptr(), dev(prv).data_ptr(), out.data_ptr(), chi, d, chi, 4 * d,
                                 wp, wb, stream()), "gauge_left")
        Q = host(Ad).reshape(chi, d * chi)
        assert np.abs(Q @ Q.T - np.eye(chi)).max() < 1e-13
        before = prv.reshape(-1, chi) @ A.reshape(chi, -1)
        after = host(out).reshape(-1, chi) @ Q
    assert relerr(after, before) < 1e-14


@pytest.mark.parametrize("chil, d, chir", [(1, 2, 2), (4, 2, 8), (32, 2, 32), (16, 2, 32), (60, 2, 60), (100, 2, 50)])
def test_gauge_moves(gpu, chil, d, chir):
    rng = np.random.default_rng(chil * chir)
    A = rng.standard_normal((chil, d, chir))
    nxt = rng.standard_normal((chir, d, 3))
    prv = rng.standard_normal((5, d, chil))
    gp, wp, wb = ws(L().tnx_gauge_workspace_bytes(chil, d, chir))
    if chil * d >= chir:
        Ad, out = dev(A), torch.empty(chir, d, 3, dtype=torch.float64, device="cuda")
        check(L().tnx_gauge_right(Ad.data_ptr(), dev(nxt).data_ptr(), out.data_ptr(), chil, d, chir, d * 3,
                                  wp, wb, stream()), "gauge_right")
        Q = host(Ad).reshape(chil * d, chir)
        assert np.linalg.norm(Q.T @ Q - np.eye(chir)) < 1e-12
        before = A.reshape(chil * d, chir) @ nxt.reshape(chir, -1)
        after = Q @ host(out).reshape(chir, -1)
        assert relerr(after, before) < 1e-13
    if d * chir >= chil:
        Ad, out = dev(A), torch.empty(5, d, chil, dtype=torch.float64, device="cuda")
        check(L().tnx_gauge_left(Ad.data_ptr(), dev(prv).data_ptr(), out.data_ptr(), chil, d, chir, 5 * d,
                                 wp, wb, stream()), "gauge_left")
        Q = host(Ad).reshape(chil, d * chir)
        assert np.linalg.norm(Q @ Q.T - np.eye(chil)) < 1e-12
        before = prv.reshape(-1, chil) @ A.reshape(chil, -1)
        after = host(out).reshape(-1, chil) @ Q
        assert relerr(after, before) < 1e-13


def test_blas1(gpu):
    rng = np.random.default_rng(0)
    x, y = rng.standard_normal(100003), rng.standard_normal(100003)
    xd, yd = dev(x), dev(y)
    check(L().tnx_axpy(x.size, 0.25, xd.data_ptr(), yd.data_ptr(), stream()))
    np.testing.assert_allclose(host(yd), y + 0.25 * x, rtol=0, atol=1e-15)
    check(L().tnx_scal(x.size, -2.0, xd.data_ptr(), stream()))
    np.testing.assert_array_equal(host(xd), -2.0 * x)
    out = C.c_double()
    _, wp, wb = ws(L().tnx_dot_workspace_bytes(x.size))
    check(L().tnx_dot(x.size, xd.data_ptr(), yd.data_ptr(), C.byref(out), wp, wb, stream()))
    assert abs(out.value - (-2.0 * x) @ (y + 0.25 * x)) < 1e-10
    # deterministic: same bits on repeat
    out2 = C.c_double()
    check(L().tnx_dot(x.size, xd.data_ptr(), yd.data_ptr(), C.byref(out2), wp, wb, stream()))
    assert out.value == out2.value


def clustered_env():
    """Ferromagnetic XXZ(n=16) after two oracle sweeps and a partial third:
    H_eff at site 8 has a cluster of low eigenvalues ~4e-4 apart on a
    spread of ~10 (the local problem of config-3-like chains)."""
    n, chi, site = 16, 16, 8
    mpo = o.xxz_mpo(n, 1.0)
    d = o.OracleDMRG(mpo, o.random_mps(n, chi, 2, seed=3))
    d.sweep(o.RIGHTWARD, 1e-10)
    d.sweep(o.LEFTWARD, 1e-10)
    for s in range(site):
        e, psi = d.one_site_solver(s, 1e-10)
        d.mps[s] = psi.reshape(d.mps[s].shape)
        d.perturb(s)
        o.split_tensor_svd(d.mps, s, o.RIGHTWARD)
        d.update_left(s + 1)
    return d.left[site], mpo[site], d.right[site], d.mps[site]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("max_basis", [10, 32])
def test_eigsh_clustered(gpu, prec, max_basis):
    from tnpy_amd import _lib

    print(f"eigsh clustered {prec} m={max_basis}", flush=True)
    Lt, W, Rt, A = clustered_env()
    M = o.heff_dense(Lt, W, Rt)
    w = np.linalg.eigvalsh(M)
    assert w[1] - w[0] < 1e-3  # the cluster is there
    shape = (Lt.shape[0], W.shape[0], W.shape[1], W.shape[2], Rt.shape[0])
    N = A.size
    vec = torch.empty(N, dtype=torch.float64, device="cuda")
    hvec = torch.empty(N, dtype=torch.float64, device="cuda")
    opts = _lib.EigshOpts(tol=1e-10, max_basis=max_basis, max_matvecs=4000, precision=PREC[prec])
    info = _lib.EigshInfo()
    _, wp, wb = ws(L().tnx_eigsh_heff1_workspace_bytes(*shape, max_basis))
    check(L().tnx_eigsh_heff1(dev(Lt).data_ptr(), dev(W).data_ptr(), dev(Rt).data_ptr(), dev(A).data_ptr(),
                              vec.data_ptr(), hvec.data_ptr(), *shape, C.byref(opts), C.byref(info), wp, wb,
                              stream()), "eigsh")
    nrm = np.abs(w).max()
    print(f"  matvecs={info.matvecs} f64={info.matvecs_f64} restarts={info.restarts}", flush=True)
    assert info.converged == 1 and info.restarts >= 1
    assert abs(info.eval - w[0]) <= 1e-12 * nrm
    v = host(vec)
    assert np.linalg.norm(M @ v - w[0] * v) <= 2e-10 * nrm


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("shape", [(1, 1, 5, 5, 2, 2), (4, 5, 5, 5, 2, 8), (16, 5, 5, 1, 2, 1),
                                   (12, 5, 5, 5, 2, 9), (7, 4, 3, 5, 3, 6), (32, 5, 5, 5, 2, 32)])
def test_heff2_matvec(gpu, prec, shape):
    chil, Dl, Dm, Dr, d, chir = shape
    print(f"heff2 {prec} {shape}", flush=True)
    rng = np.random.default_rng(sum(shape) + 7)
    Lt = rng.standard_normal((chil, Dl, chil))
    W1 = rng.standard_normal((Dl, Dm, d, d))
    W2 = rng.standard_normal((Dm, Dr, d, d))
    Rt = rng.standard_normal((chir, Dr, chir))
    x = rng.standard_normal((chil, d, d, chir))
    y = torch.empty(x.size, dtype=torch.float64, device="cuda")
    _, wp, wb = ws(L().tnx_heff2_workspace_bytes(*shape))
    check(L().tnx_heff2_matvec(dev(Lt).data_ptr(), dev(W1).data_ptr(), dev(W2).data_ptr(), dev(Rt).data_ptr(),
                               dev(x).data_ptr(), y.data_ptr(), *shape, PREC[prec], wp, wb, stream()), "heff2")
    ref = o.two_site_matvec(Lt, W1, W2, Rt, x)
    assert relerr(host(y), ref) < 10 * TOL[prec]


# ---------------------------------------------------------------------------
def _spectrum_matrix(rng, a, b, kind):
    n = min(a, b)
    if kind == "random":
        return rng.standard_normal((a, b))
    U, _ = np.linalg.qr(rng.standard_normal((a, n)))
    V, _ = np.linalg.qr(rng.standard_normal((b, n)))
    if kind == "graded":
        s = np.logspace(0, -12, n)
    elif kind == "rank":
        s = np.concatenate([np.linspace(2.0, 0.5, n // 5), np.zeros(n - n // 5)])
    else:  # degenerate: Schmidt-like multiplets
        s = np.repeat(np.logspace(0, -6, max(1, n // 4)), 4)[:n]
        s = np.concatenate([s, np.zeros(n - len(s))])
    return (U * s) @ V.T


@pytest.mark.parametrize("a, b, k, kind", [(1, 1, 1, "random"), (7, 5, 3, "random"), (37, 23, 10, "random"),
                                           (23, 37, 10, "random"), (130, 130, 64, "graded"),
                                           (256, 200, 200, "rank"), (300, 512, 150, "degenerate"),
                                           (512, 512, 256, "graded")])
def test_svd_trunc(gpu, a, b, k, kind):
    """Top-k singular triplets against numpy.linalg.svd: singular values to
    1e-12 * s_max, U orthonormal, Y = U^T X, and the truncation error equal
    to the discarded weight (Eckart-Young)."""
    from tnpy_amd import _lib

    print(f"svd_trunc {a}x{b} k={k} {kind}", flush=True)
    rng = np.random.default_rng(a * 7 + b)
    X = _spectrum_matrix(rng, a, b, kind)
    U, Y, sv = _lib.svd_trunc(dev(X), k)
    s = np.linalg.svd(X, compute_uv=False)
    smax = max(s[0], 1e-300)
    Uh, Yh = host(U), host(Y)
    assert np.abs(sv - s[:k]).max() <= 1e-12 * smax, np.abs(sv - s[:k]).max() / smax
    assert np.abs(Uh.T @ Uh - np.eye(k)).max() < 1e-12
    assert np.abs(Yh - Uh.T @ X).max() <= 1e-12 * smax
    resid = np.linalg.norm(X - Uh @ Yh)
    best = np.sqrt(np.sum(s[k:] ** 2))
    assert abs(resid - best) <= 1e-10 * smax + 1e-8 * best, (resid, best)
