"""Batched entry points and BatchedFiniteDMRG.

Every batched call is checked chain by chain against the CPU oracle and
against the single-chain entry point on the same chain (FP64, ~1e-13
relative for contractions and QR; eigenvalues to 1e-12 relative). The
driver is checked against FiniteDMRG runs from the same initial states and
against exact diagonalisation at small n.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import dmrg_ref as o

pytestmark = pytest.mark.gpu

_KEEP = []


def dev(a):
    t = torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")
    _KEEP.append(t)
    return t


@pytest.fixture(autouse=True)
def _release():
    yield
    torch.cuda.synchronize()
    _KEEP.clear()


def lib():
    from tnpy_amd import _lib

    return _lib.lib()


def ws(nbytes):
    buf = torch.empty(int(nbytes) + 256, dtype=torch.uint8, device="cuda")
    _KEEP.append(buf)
    return buf.data_ptr(), buf.numel()


def stream():
    return torch.cuda.current_stream().cuda_stream


def check(st, what="call"):
    assert st == 0, f"{what}: status {st}: {lib().tnx_last_error()}"


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def sym_env(rng, chi, D):
    """Random environment, symmetric under exchanging ket and bra legs
    (as every environment of a real MPS is)."""
    E = rng.standard_normal((chi, D, chi))
    return 0.5 * (E + E.transpose(2, 1, 0))


def mpo_stack(B, D, d, rng, symmetric=False):
    """B random-coupling Heisenberg bulk W tensors (D=5, d=2). With
    symmetric=True each W[w][v] is symmetrised in (ket, bra), so that with
    symmetric environments H_eff is a symmetric matrix (an eigensolver
    input); the physical W is symmetric only after summing the S+ and S-
    channels, which random environments do not pair up."""
    Ws = []
    for b in range(B):
        J = rng.uniform(0, 1, size=2)
        W = o.random_coupling_heisenberg_mpo(3, J)[1]
        Ws.append(0.5 * (W + W.transpose(0, 1, 3, 2)) if symmetric else W)
    return np.stack(Ws)


SHAPES = [(4, 8, 3), (16, 16, 5), (64, 64, 7)]  # (chil, chir, batch)


@pytest.mark.parametrize("chil, chir, B", SHAPES)
def test_matvec_batched(gpu, chil, chir, B):
    print(f"matvec_batched chil={chil} chir={chir} B={B}", flush=True)
    rng = np.random.default_rng(chil + chir + B)
    D, d = 5, 2
    Ls = np.stack([sym_env(rng, chil, D) for _ in range(B)])
    Rs = np.stack([sym_env(rng, chir, D) for _ in range(B)])
    Ws = mpo_stack(B, D, d, rng)
    xs = rng.standard_normal((B, chil * d * chir))
    y = torch.empty(B, chil * d * chir, dtype=torch.float64, device="cuda")
    wp, wb = ws(lib().tnx_heff1_batched_workspace_bytes(chil, D, D, d, chir, B))
    check(lib().tnx_heff1_matvec_batched(dev(Ls).data_ptr(), dev(Ws).data_ptr(), dev(Rs).data_ptr(),
                                         dev(xs).data_ptr(), y.data_ptr(), chil, D, D, d, chir, B, 0,
                                         wp, wb, stream()))
    yb = host(y)
    ys = torch.empty(chil * d * chir, dtype=torch.float64, device="cuda")
    wp1, wb1 = ws(lib().tnx_heff1_workspace_bytes(chil, D, D, d, chir))
    for b in range(B):
        ref = o.heff_matvec(Ls[b], Ws[b], Rs[b], xs[b])
        assert relerr(yb[b], ref) < 1e-13
        check(lib().tnx_heff1_matvec(dev(Ls[b]).data_ptr(), dev(Ws[b]).data_ptr(), dev(Rs[b]).data_ptr(),
                                     dev(xs[b]).data_ptr(), ys.data_ptr(), chil, D, D, d, chir, 0, wp1, wb1,
                                     stream()))
        assert relerr(yb[b], host(ys)) < 1e-14


@pytest.mark.parametrize("chil, chir, B", SHAPES)
def test_env_batched(gpu, chil, chir, B):
    print(f"env_batched chil={chil} chir={chir} B={B}", flush=True)
    rng = np.random.default_rng(3 * chil + chir + B)
    D, d = 5, 2
    Ls = np.stack([sym_env(rng, chil, D) for _ in range(B)])
    Rs = np.stack([sym_env(rng, chir, D) for _ in range(B)])
    Ws = mpo_stack(B, D, d, rng)
    As = rng.standard_normal((B, chil, d, chir))
    Lo = torch.empty(B, chir, D, chir, dtype=torch.float64, device="cuda")
    Ro = torch.empty(B, chil, D, chil, dtype=torch.float64, device="cuda")
    wp, wb = ws(lib().tnx_env_batched_workspace_bytes(chil, D, D, d, chir, B))
    check(lib().tnx_env_left_batched(dev(Ls).data_ptr(), dev(As).data_ptr(), dev(Ws).data_ptr(),
                                     Lo.data_ptr(), chil, D, D, d, chir, B, 0, wp, wb, stream()))
    check(lib().tnx_env_right_batched(dev(Rs).data_ptr(), dev(As).data_ptr(), dev(Ws).data_ptr(),
                                      Ro.data_ptr(), chil, D, D, d, chir, B, 0, wp, wb, stream()))
    lo, ro = host(Lo), host(Ro)
    for b in range(B):
        assert relerr(lo[b], o.env_left(Ls[b], As[b], Ws[b])) < 1e-13
        assert relerr(ro[b], o.env_right(Rs[b], As[b], Ws[b])) < 1e-13


@pytest.mark.parametrize("chil, chir, B", [(2, 4, 5), (4, 8, 9), (8, 8, 3)])
def test_dense_and_syev_batched(gpu, chil, chir, B):
    print(f"dense_syev_batched chil={chil} chir={chir} B={B}", flush=True)
    rng = np.random.default_rng(chil * chir + B)
    D, d = 5, 2
    N = chil * d * chir
    Ls = np.stack([sym_env(rng, chil, D) for _ in range(B)])
    Rs = np.stack([sym_env(rng, chir, D) for _ in range(B)])
    Ws = mpo_stack(B, D, d, rng, symmetric=True)
    M = torch.empty(B, N, N, dtype=torch.float64, device="cuda")
    check(lib().tnx_heff1_dense_batched(dev(Ls).data_ptr(), dev(Ws).data_ptr(), dev(Rs).data_ptr(),
                                        M.data_ptr(), chil, D, D, d, chir, B, stream()))
    Mh = host(M)
    evec = torch.empty(B, N, dtype=torch.float64, device="cuda")
    ev = (C.c_double * B)()
    wp, wb = ws(lib().tnx_syev_lowest_batched_workspace_bytes(N, B))
    check(lib().tnx_syev_lowest_batched(M.data_ptr(), N, B, evec.data_ptr(), ev, wp, wb, stream()))
    vh = host(evec)
    for b in range(B):
        ref = o.heff_dense(Ls[b], Ws[b], Rs[b])
        assert relerr(Mh[b], ref) < 1e-13
        assert np.abs(ref - ref.T).max() < 1e-12 * np.abs(ref).max()
        w, V = np.linalg.eigh(ref)
        assert abs(ev[b] - w[0]) <= 1e-12 * max(1.0, abs(w).max())
        # eigenvector up to sign (gap permitting)
        if w[1] - w[0] > 1e-6 * abs(w).max():
            assert min(np.linalg.norm(vh[b] - V[:, 0]), np.linalg.norm(vh[b] + V[:, 0])) < 1e-8


def test_syev_batched_n1(gpu):
    B = 4
    M = dev(np.arange(1.0, B + 1.0).reshape(B, 1, 1))
    evec = torch.empty(B, 1, dtype=torch.float64, device="cuda")
    ev = (C.c_double * B)()
    wp, wb = ws(lib().tnx_syev_lowest_batched_workspace_bytes(1, B))
    check(lib().tnx_syev_lowest_batched(M.data_ptr(), 1, B, evec.data_ptr(), ev, wp, wb, stream()))
    assert list(ev) == [1.0, 2.0, 3.0, 4.0]
    assert np.all(host(evec) == 1.0)


@pytest.mark.parametrize("steps_per_check", [1, 8])
@pytest.mark.parametrize("chil, chir, B", [(2, 4, 3), (16, 16, 4), (32, 64, 6), (64, 64, 5)])
def test_eigsh_batched(gpu, chil, chir, B, steps_per_check):
    """Every chain's pair against the oracle matvec and the single-chain
    solver (same start: same eigenvalue and matvec count), with statuses
    read after every step and with convergence decided on the device over
    runs of 8 steps; (2, 4) exhausts its 16-dimensional space."""
    print(f"eigsh_batched chil={chil} chir={chir} B={B} steps_per_check={steps_per_check}", flush=True)
    from tnpy_amd import _lib

    rng = np.random.default_rng(chil + 5 * chir + B)
    D, d = 5, 2
    N = chil * d * chir
    Ls = np.stack([sym_env(rng, chil, D) for _ in range(B)])
    Rs = np.stack([sym_env(rng, chir, D) for _ in range(B)])
    Ws = mpo_stack(B, D, d, rng, symmetric=True)
    v0 = rng.standard_normal((B, N))
    vec = torch.empty(B, N, dtype=torch.float64, device="cuda")
    hvec = torch.empty(B, N, dtype=torch.float64, device="cuda")
    opts = _lib.EigshOpts(1e-10, 32, 4000, 0, 8, steps_per_check)
    infos = (_lib.EigshInfo * B)()
    wp, wb = ws(lib().tnx_eigsh_heff1_batched_workspace_bytes(chil, D, D, d, chir, 32, B))
    check(lib().tnx_eigsh_heff1_batched(dev(Ls).data_ptr(), dev(Ws).data_ptr(), dev(Rs).data_ptr(),
                                        dev(v0).data_ptr(), vec.data_ptr(), hvec.data_ptr(), chil, D, D, d,
                                        chir, B, C.byref(opts), infos, wp, wb, stream()), "eigsh_batched")
    vb, hb = host(vec), host(hvec)
    s_vec = torch.empty(N, dtype=torch.float64, device="cuda")
    s_hvec = torch.empty(N, dtype=torch.float64, device="cuda")
    wp1, wb1 = ws(lib().tnx_eigsh_heff1_workspace_bytes(chil, D, D, d, chir, 32))
    counts = set()
    for b in range(B):
        info = infos[b]
        assert info.converged == 1
        counts.add(info.matvecs)
        # the pair satisfies H y = theta y to the criterion, in FP64
        hy = o.heff_matvec(Ls[b], Ws[b], Rs[b], vb[b])
        assert abs(np.linalg.norm(vb[b]) - 1.0) < 1e-12
        assert np.linalg.norm(hy - hb[b]) <= 1e-12 * np.linalg.norm(hy) + 1e-14
        assert np.linalg.norm(hy - info.eval * vb[b]) <= 2e-10 * info.anorm
        # same answer as the single-chain solver from the same start
        si = _lib.EigshInfo()
        check(lib().tnx_eigsh_heff1(dev(Ls[b]).data_ptr(), dev(Ws[b]).data_ptr(), dev(Rs[b]).data_ptr(),
                                    dev(v0[b]).data_ptr(), s_vec.data_ptr(), s_hvec.data_ptr(), chil, D, D, d,
                                    chir, C.byref(opts), C.byref(si), wp1, wb1, stream()), "eigsh")
        assert abs(si.eval - info.eval) <= 1e-12 * info.anorm
        assert si.matvecs == info.matvecs
        if N <= 2048:
            w = np.linalg.eigvalsh(o.heff_dense(Ls[b], Ws[b], Rs[b]))
            assert abs(info.eval - w[0]) <= 1e-9 * abs(w).max()
    print(f"matvec counts per chain: {sorted(counts)}")


@pytest.mark.parametrize("steps_per_check", [1, 8])
def test_eigsh_batched_matvec_limit(gpu, steps_per_check):
    """A matvec limit below what the chains need: TNX_ERR_NOCONV, every
    chain flagged unconverged, and the returned pair still consistent (unit
    vector, hvec = H vec, eval its Rayleigh quotient)."""
    from tnpy_amd import _lib

    rng = np.random.default_rng(77)
    chil = chir = 16
    B, D, d = 3, 5, 2
    N = chil * d * chir
    Ls = np.stack([sym_env(rng, chil, D) for _ in range(B)])
    Rs = np.stack([sym_env(rng, chir, D) for _ in range(B)])
    Ws = mpo_stack(B, D, d, rng, symmetric=True)
    v0 = rng.standard_normal((B, N))
    vec = torch.empty(B, N, dtype=torch.float64, device="cuda")
    hvec = torch.empty(B, N, dtype=torch.float64, device="cuda")
    opts = _lib.EigshOpts(1e-14, 32, 10, 0, 8, steps_per_check)
    infos = (_lib.EigshInfo * B)()
    wp, wb = ws(lib().tnx_eigsh_heff1_batched_workspace_bytes(chil, D, D, d, chir, 32, B))
    st = lib().tnx_eigsh_heff1_batched(dev(Ls).data_ptr(), dev(Ws).data_ptr(), dev(Rs).data_ptr(),
                                       dev(v0).data_ptr(), vec.data_ptr(), hvec.data_ptr(), chil, D, D, d,
                                       chir, B, C.byref(opts), infos, wp, wb, stream())
    assert st == -3
    vb, hb = host(vec), host(hvec)
    for b in range(B):
        info = infos[b]
        assert info.converged == 0 and info.matvecs == 10
        hy = o.heff_matvec(Ls[b], Ws[b], Rs[b], vb[b])
        assert abs(np.linalg.norm(vb[b]) - 1.0) < 1e-12
        assert np.linalg.norm(hy - hb[b]) <= 1e-11 * np.linalg.norm(hy)
        assert abs(vb[b] @ hy - info.eval) <= 1e-11 * info.anorm


def test_eigsh_batched_rejects_low_precision(gpu):
    from tnpy_amd import _lib

    opts = _lib.EigshOpts(1e-10, 32, 100, 1, 8)
    infos = (_lib.EigshInfo * 2)()
    x = dev(np.ones(64))
    st = lib().tnx_eigsh_heff1_batched(x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(),
                                       x.data_ptr(), 1, 1, 1, 2, 1, 2, C.byref(opts), infos, x.data_ptr(),
                                       1 << 30, stream())
    assert st == -4


@pytest.mark.parametrize("chil, d, chir, B", [(1, 2, 2, 3), (8, 2, 16, 4), (64, 2, 64, 5), (32, 2, 64, 2)])
def test_gauge_batched(gpu, chil, d, chir, B):
    print(f"gauge_batched {chil}x{d}x{chir} B={B}", flush=True)
    rng = np.random.default_rng(chil * d + chir + B)
    # rightward: A (chil*d x chir), next (chir x 7)
    A = rng.standard_normal((B, chil, d, chir))
    nxt = rng.standard_normal((B, chir, 7))
    Ad, nd = dev(A), dev(nxt)
    out = torch.empty(B, chir, 7, dtype=torch.float64, device="cuda")
    wp, wb = ws(lib().tnx_gauge_batched_workspace_bytes(chil, d, chir, B))
    check(lib().tnx_gauge_right_batched(Ad.data_ptr(), nd.data_ptr(), out.data_ptr(), chil, d, chir, 7, B,
                                        wp, wb, stream()))
    Q, no = host(Ad), host(out)
    wp1, wb1 = ws(lib().tnx_gauge_workspace_bytes(chil, d, chir))
    for b in range(B):
        q = Q[b].reshape(chil * d, chir)
        assert np.abs(q.T @ q - np.eye(chir)).max() < 1e-13
        # state unchanged: A next == Q (R next)
        assert relerr(q @ no[b], A[b].reshape(chil * d, chir) @ nxt[b]) < 1e-13
        a1, n1 = dev(A[b]), dev(nxt[b])
        o1 = torch.empty(chir, 7, dtype=torch.float64, device="cuda")
        check(lib().tnx_gauge_right(a1.data_ptr(), n1.data_ptr(), o1.data_ptr(), chil, d, chir, 7, wp1, wb1,
                                    stream()))
        assert relerr(Q[b], host(a1)) < 1e-12
        assert relerr(no[b], host(o1)) < 1e-12
    # leftward: A (chil x d*chir) with chil <= d*chir; prev (5 x chil)
    if chil <= d * chir:
        prv = rng.standard_normal((B, 5, chil))
        Ad, pd = dev(A), dev(prv)
        out = torch.empty(B, 5, chil, dtype=torch.float64, device="cuda")
        check(lib().tnx_gauge_left_batched(Ad.data_ptr(), pd.data_ptr(), out.data_ptr(), chil, d, chir, 5, B,
                                           wp, wb, stream()))
        Q, po = host(Ad), host(out)
        for b in range(B):
            q = Q[b].reshape(chil, d * chir)
            assert np.abs(q @ q.T - np.eye(chil)).max() < 1e-13
            assert relerr(po[b] @ q, prv[b] @ A[b].reshape(chil, d * chir)) < 1e-13


def test_gauge_batched_rejects_large(gpu):
    x = dev(np.ones(8))
    wp, wb = ws(lib().tnx_gauge_batched_workspace_bytes(256, 2, 256, 1))
    st = lib().tnx_gauge_right_batched(x.data_ptr(), x.data_ptr() + 8, x.data_ptr() + 16, 256, 2, 256, 1, 1,
                                       wp, wb, stream())
    assert st == -1


# ---------------------------------------------------------------------------
def _chains(ns, seed):
    from tnpy_amd.model import RandomCouplingHeisenberg

    rng = np.random.default_rng(seed)
    return [RandomCouplingHeisenberg(n=n, seed=int(rng.integers(1 << 30))) for n in ns]


@pytest.mark.parametrize("mode", ["grouped", "serial", "concurrent"])
def test_batched_dmrg_matches_single(gpu, mode):
    """Per chain, sweep by sweep, the batched driver equals FiniteDMRG from
    the same initial state (mixed lengths, dense and Krylov sites): length
    groups stepped in lockstep with shared launches, swept one after
    another, and swept concurrently."""
    from tnpy_amd import Direction, FiniteDMRG, MatrixProductState
    from tnpy_amd.batched_dmrg import BatchedFiniteDMRG

    ns = [10, 14, 10, 14, 12]
    chi = 16
    models = _chains(ns, 7)
    starts = [MatrixProductState.random(n, chi, 2, seed=100 + i, device=gpu) for i, n in enumerate(ns)]
    bd = BatchedFiniteDMRG([m.mpo for m in models], bond_dim=chi,
                           mps=[s.copy() for s in starts], device=gpu, grouped=mode == "grouped",
                           concurrent=mode == "concurrent")
    assert bd.groups == {10: [0, 2], 12: [4], 14: [1, 3]}
    dirs = [Direction.RIGHTWARD, Direction.LEFTWARD, Direction.RIGHTWARD]
    eb = np.array([bd.sweep(dr, tol=1e-12) for dr in dirs]).T
    solves = bd.stats["dense_sites"] + bd.stats["krylov_sites"]
    per_group = len(dirs) * sum(n - 1 for n in (10, 12, 14))
    print(f"{mode}: {solves} solve launches for {per_group} group site steps", flush=True)
    if mode == "grouped":
        # at least one step per site of the longest chain, and fewer solve
        # launches than group site steps (shared stacks)
        assert len(dirs) * 13 <= solves < per_group
    else:
        assert solves == per_group
    for i, m in enumerate(models):
        f = FiniteDMRG(m.mpo, bond_dim=chi, mps=starts[i].copy(), compute_variance=False, device=gpu)
        es = [f.sweep(dr, tol=1e-12) for dr in dirs]
        np.testing.assert_allclose(eb[i], es, rtol=1e-11, atol=0)


def test_batched_dmrg_vs_ed(gpu):
    """Random-coupling Heisenberg chains, n=8 and 10: ground energies from
    the batched driver against exact diagonalisation of each chain."""
    from tnpy_amd.batched_dmrg import BatchedFiniteDMRG

    ns = [8, 10, 8, 10, 10, 8]
    models = _chains(ns, 11)
    bd = BatchedFiniteDMRG([m.mpo for m in models], bond_dim=32, seeds=list(range(len(ns))), device=gpu)
    eb = bd.run(tol=1e-12, max_sweep=12)
    for i, m in enumerate(models):
        H = o.full_hamiltonian(o.random_coupling_heisenberg_mpo(ns[i], m.couplings))
        e0 = np.linalg.eigvalsh(H)[0]
        assert abs(eb[i][-1] - e0) <= 1e-10 * abs(e0), (i, eb[i][-1], e0)


def test_batched_dmrg_bulk_krylov(gpu):
    """chi=64 bulk sites (N=8192, Krylov) for two lengths, 3 sweeps, against
    FiniteDMRG on two of the chains."""
    from tnpy_amd import Direction, FiniteDMRG, MatrixProductState
    from tnpy_amd.batched_dmrg import BatchedFiniteDMRG

    ns = [16, 20, 16, 20]
    chi = 64
    models = _chains(ns, 3)
    starts = [MatrixProductState.random(n, chi, 2, seed=40 + i, device=gpu) for i, n in enumerate(ns)]
    bd = BatchedFiniteDMRG([m.mpo for m in models], chi=chi, mps=[s.copy() for s in starts], device=gpu)
    dirs = [Direction.RIGHTWARD, Direction.LEFTWARD, Direction.RIGHTWARD]
    eb = np.array([bd.sweep(dr, tol=1e-10) for dr in dirs]).T
    assert bd.stats["krylov_sites"] > 0
    for i in (0, 1):
        f = FiniteDMRG(models[i].mpo, bond_dim=chi, mps=starts[i].copy(), compute_variance=False, device=gpu)
        es = [f.sweep(dr, tol=1e-10) for dr in dirs]
        np.testing.assert_allclose(eb[i], es, rtol=1e-10, atol=0)
    # the state of a chain comes back as an MPS with the driver's energy
    from tnpy_amd import Environment

    st = bd.mps(1)
    env = Environment(models[1].mpo, st)
    e = env.expectation() / st.norm() ** 2
    assert abs(e - eb[1][-1]) <= 1e-8 * abs(eb[1][-1])
    # run() keeps each chain's energies up to its own convergence
    hist = bd.run(tol=1e-9, max_sweep=4)
    assert all(1 <= len(h) <= 4 for h in hist)


def test_batched_random_states_match_single(gpu):
    """The batched initial states equal MatrixProductState.random with the
    same seeds (same draws, LQ sweep and rescaling) to rounding."""
    from tnpy_amd import MatrixProductState
    from tnpy_amd.batched_dmrg import BatchedFiniteDMRG

    ns = [10, 12, 10]
    seeds = [5, 9, 13]
    models = _chains(ns, 1)
    bd = BatchedFiniteDMRG([m.mpo for m in models], bond_dim=16, seeds=seeds, device=gpu)
    for i, n in enumerate(ns):
        ref = MatrixProductState.random(n, 16, 2, seed=seeds[i], device=gpu, host_rng=True)
        got = bd.mps(i)
        for a, b in zip(got.tensors, ref.tensors):
            assert a.shape == b.shape
            assert (a - b).abs().max().item() <= 1e-12 * max(1.0, b.abs().max().item())
