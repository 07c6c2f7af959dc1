"""The CPU oracle pinned against the reference's own literals, the
notebook's printed ED spectrum, closed forms and brute force."""
import numpy as np
import pytest

from oracle import dmrg_ref as o


class TestOperatorsLiterals:
    """tests/test_operators.py:39-98 of the reference (exact equality)."""

    @pytest.mark.parametrize(
        "key, mpo",
        [
            ("random_heisenberg_n2_h0", lambda: o.random_heisenberg_mpo(2, 0.0)),
            ("random_heisenberg_n3_h0", lambda: o.random_heisenberg_mpo(3, 0.0)),
            ("xxz_n2_delta0.5", lambda: o.xxz_mpo(2, 0.5)),
            ("xxz_n3_delta0.5", lambda: o.xxz_mpo(3, 0.5)),
        ],
    )
    def test_full_hamiltonian(self, fixtures, key, mpo):
        np.testing.assert_array_equal(o.full_hamiltonian(mpo()), np.array(fixtures["operators"][key]))

    def test_spin_half(self):
        Sp, Sm, Sz, I2, O2 = o.spin_operators()
        np.testing.assert_array_equal(Sp + Sm, [[0, 1], [1, 0]])
        np.testing.assert_array_equal(-1j * (Sp - Sm), [[0, -1j], [1j, 0]])

    @pytest.mark.parametrize("h", [0.0, 0.5])
    def test_square(self, h):
        # tests/test_operators.py:19-28
        mpo = o.random_heisenberg_mpo(4, h, seed=7)
        sq = o.mpo_square(mpo)
        assert sq[0].shape == (1, 25, 2, 2) and sq[1].shape == (25, 25, 2, 2) and sq[3].shape == (25, 1, 2, 2)
        H = o.full_hamiltonian(mpo)
        np.testing.assert_allclose(H @ H, o.full_hamiltonian(sq), atol=1e-12)

    @pytest.mark.parametrize("offset", [0.5, -0.1])
    def test_offset(self, offset):
        # tests/model/test_random_heisenberg.py:10-26
        a = o.full_hamiltonian(o.random_heisenberg_mpo(6, 0.5, seed=2022))
        b = o.full_hamiltonian(o.random_heisenberg_mpo(6, 0.5, seed=2022, offset=offset))
        np.testing.assert_allclose(a - offset * np.eye(64), b, atol=1e-12)

    def test_penalty_is_total_sz_squared(self):
        # the penalty row/column of random_heisenberg.py:50-58,84-109:
        # H(lambda) - H(0) = lambda (sum Sz - S_target)^2
        n, lam, st = 5, 0.7, 1
        H0 = o.full_hamiltonian(o.random_heisenberg_mpo(n, 0.3, seed=3))
        Hp = o.full_hamiltonian(o.random_heisenberg_mpo(n, 0.3, penalty=lam, s_target=st, seed=3))
        sz = np.array([0.5 * (n - 2 * bin(k).count("1")) for k in range(2**n)])
        np.testing.assert_allclose(Hp - H0, lam * np.diag((sz - st) ** 2), atol=1e-12)


class TestED:
    def test_heisenberg_n2(self, fixtures):
        # tests/test_exact_diagonalization.py:15-42
        ed = o.ExactDiagonalization(o.random_heisenberg_mpo(2, 0.0))
        np.testing.assert_array_equal(ed.matrix, fixtures["ed"]["heisenberg_n2_matrix"])
        np.testing.assert_allclose(ed.evals, fixtures["ed"]["heisenberg_n2_evals"], atol=1e-15)
        g = np.array(fixtures["ed"]["heisenberg_n2_evecs"])[:, 0]
        assert abs(abs(g @ ed.evecs[:, 0]) - 1) < 1e-14

    def test_notebook_spectrum(self, fixtures):
        # docs/source/notebooks/tsdrg.ipynb:140-150, printed to 8 digits
        ed = o.ExactDiagonalization(o.random_heisenberg_mpo(6, 3.0, seed=2021))
        ref = np.array(fixtures["notebook_random_heisenberg_n6_h3_seed2021_evals"])
        np.testing.assert_allclose(ed.evals, ref, atol=6e-8)

    def test_fields(self, fixtures):
        np.testing.assert_array_equal(o.random_fields(6, 3.0, 2021), fixtures["random_fields"]["n6_h3_seed2021"])

    @pytest.mark.parametrize("n", [10, 16])
    def test_xx_closed_form(self, fixtures, n):
        # delta = 0: free fermions; sparse ED of the MPO-defined Hamiltonian
        import scipy.sparse.linalg as spla

        e = spla.eigsh(o.xxz_sparse(n, 0.0), k=1, which="SA", tol=1e-14)[0][0]
        assert abs(e - fixtures["closed_form"]["xx_chain_E0"][str(n)]) < 1e-11

    def test_ferro_closed_form(self):
        ed = o.ExactDiagonalization(o.xxz_mpo(10, 1.0))
        assert abs(ed.evals[0] - (-(10 - 1) / 4)) < 1e-12

    def test_sparse_matches_mpo(self):
        np.testing.assert_allclose(o.xxz_sparse(8, 0.7).toarray(), o.full_hamiltonian(o.xxz_mpo(8, 0.7)), atol=1e-14)

    def test_random_coupling_model(self):
        # H = sum J_i S_i.S_{i+1}, brute force from Kronecker products
        n = 6
        J = np.random.default_rng(1).uniform(0, 1, n - 1)
        Sp, Sm, Sz, _, _ = o.spin_operators()
        H = np.zeros((2**n, 2**n))
        for i in range(n - 1):
            for c, A, B in ((0.5, Sp, Sm), (0.5, Sm, Sp), (1.0, Sz, Sz)):
                H += J[i] * c * np.kron(np.kron(np.eye(2**i), np.kron(A, B)), np.eye(2 ** (n - i - 2)))
        np.testing.assert_allclose(o.full_hamiltonian(o.random_coupling_heisenberg_mpo(n, J)), H, atol=1e-14)


class TestOracleDMRG:
    def test_bond_dims(self):
        # tests/test_matrix_product_state.py:11-37
        assert o.bond_dims(8, 6, 2) == [2, 4, 6, 6, 6, 4, 2]
        assert o.bond_dims(6, 4, 4) == [4, 4, 4, 4, 4]

    def test_random_mps_canonical(self):
        mps = o.random_mps(8, 10, 2, seed=3)
        assert abs(np.linalg.norm(o.mps_to_dense(mps)) - 1) < 1e-12
        for A in mps[1:]:
            l, p, r = A.shape
            M = A.reshape(l, p * r)
            np.testing.assert_allclose(M @ M.T, np.eye(l), atol=1e-12)

    @pytest.mark.parametrize("direction", [o.RIGHTWARD, o.LEFTWARD])
    def test_split_preserves_state(self, direction):
        # tests/test_matrix_product_state.py:83-89
        mps = o.random_mps(8, 10, 2, seed=4)
        before = o.mps_to_dense(mps)
        site = 3 if direction == o.RIGHTWARD else 4
        o.split_tensor_svd(mps, site, direction)
        np.testing.assert_allclose(o.mps_to_dense(mps), before, atol=1e-12)

    def test_matvec_matches_dense_and_ed_projection(self):
        # H_eff via the two reference code paths, and P^T H P from ED
        n = 6
        mpo = o.random_heisenberg_mpo(n, 0.5, seed=5)
        mps = o.random_mps(n, 4, 2, seed=6)
        d = o.OracleDMRG(mpo, mps)
        H = o.full_hamiltonian(mpo)
        for site in range(n):
            M = o.heff_dense(d.left[site], mpo[site], d.right[site])
            N = M.shape[0]
            cols = np.stack([d.matvec(site, e) for e in np.eye(N)], axis=1)
            np.testing.assert_allclose(cols, M.T, atol=1e-13)
            np.testing.assert_allclose(M, M.T, atol=1e-13)
            # embedding P: x (l,k,r) -> left-canonical part, site, right part
            P = np.zeros((2**n, N))
            for c in range(N):
                arr = [a.copy() for a in mps]
                arr[site] = np.eye(N)[:, c].reshape(mps[site].shape)
                P[:, c] = o.mps_to_dense(arr)
            np.testing.assert_allclose(P.T @ H @ P, M, atol=1e-11)

    def test_expectation_equals_ed(self):
        n = 6
        mpo = o.xxz_mpo(n, 0.5)
        mps = o.random_mps(n, 8, 2, seed=9)
        v = o.mps_to_dense(mps)
        assert abs(o.expectation(mps, mpo) - v @ o.full_hamiltonian(mpo) @ v) < 1e-12
        H = o.full_hamiltonian(mpo)
        var = v @ H @ H @ v - (v @ H @ v) ** 2
        assert abs(o.variance(mps, mpo) - var) < 1e-12

    def test_dmrg_xxz_n10(self, fixtures):
        # tests/test_finite_dmrg.py:9-23 (bond_dim=32, tol=1e-8, atol 1e-8)
        d = o.OracleDMRG(o.xxz_mpo(10, 0.5), o.random_mps(10, 32, 2, seed=0))
        e = d.run(tol=1e-8)[-1]
        assert abs(e - fixtures["brute_force"]["xxz_n10_delta0.5_E0"]) < 1e-8


def _embed(mps, sites, block_shape):
    """Dense embedding P: columns are the dense states with the tensors of
    ``sites`` (one site, or two adjacent) replaced by each unit block."""
    N = int(np.prod(block_shape))
    cols = []
    for c in range(N):
        blk = np.eye(N)[:, c].reshape(block_shape)
        if len(sites) == 1:
            arr = [a.copy() for a in mps]
            arr[sites[0]] = blk
        else:
            s = sites[0]
            l, k1, k2, r = block_shape
            # the block as one site of physical dim k1*k2 (ordering k1 major, as
            # the dense state fuses physical indices left to right)
            arr = [a.copy() for a in mps[:s]] + [blk.reshape(l, k1 * k2, r)] + [a.copy() for a in mps[s + 2:]]
        cols.append(o.mps_to_dense(arr))
    return np.stack(cols, axis=1)


class TestTwoSiteOracle:
    """The two-site functions have no reference counterpart (the reference's
    two-site solver is a stub, tnpy/finite_dmrg.py:113-114); they are pinned
    by P^T H P from exact diagonalisation, by ED energies, and by the
    Eckart-Young identity of the truncated split."""

    @pytest.mark.parametrize("site", [0, 2, 4])
    def test_two_site_matvec_is_projected_hamiltonian(self, site):
        n = 6
        mpo = o.random_heisenberg_mpo(n, 0.5, seed=11)
        mps = o.random_mps(n, 4, 2, seed=12)
        d = o.OracleDMRG(mpo, mps)
        L, R = d.left[site], d.right[site + 1]
        W1, W2 = mpo[site], mpo[site + 1]
        shape = (mps[site].shape[0], 2, 2, mps[site + 1].shape[2])
        N = int(np.prod(shape))
        P = _embed(mps, [site, site + 1], shape)
        H = o.full_hamiltonian(mpo)
        ref = P.T @ H @ P
        cols = np.stack([o.two_site_matvec(L, W1, W2, R, e.reshape(shape)) for e in np.eye(N)], axis=1)
        np.testing.assert_allclose(cols, ref, atol=1e-11)
        # the dense two-site H_eff of the oracle's exact branch
        M = np.einsum("awb,wukc,uvld,rvs->aklrbcds", L, W1, W2, R).reshape(N, N)
        np.testing.assert_allclose(M, ref, atol=1e-11)

    def test_one_site_projection_is_two_site_restricted(self):
        # a two-site vector theta = A . B with B fixed gives the one-site
        # H_eff of site s on the right-orthonormal B
        n = 6
        mpo = o.xxz_mpo(n, 0.5)
        mps = o.random_mps(n, 4, 2, seed=13)
        d = o.OracleDMRG(mpo, mps)
        s = 2
        x = np.random.default_rng(0).standard_normal(mps[s].shape)
        theta = np.tensordot(x, mps[s + 1], axes=(2, 0))
        y2 = o.two_site_matvec(d.left[s], mpo[s], mpo[s + 1], d.right[s + 1], theta).reshape(theta.shape)
        B = mps[s + 1]
        y1 = np.tensordot(y2, B, axes=([2, 3], [1, 2]))  # project back with B^T
        np.testing.assert_allclose(y1.reshape(-1), d.matvec(s, x.reshape(-1)), atol=1e-12)

    @pytest.mark.parametrize("n, chi, delta", [(8, 16, 0.5), (10, 32, 0.5), (8, 16, 0.0)])
    def test_dmrg2_energy_matches_ed(self, n, chi, delta):
        mpo = o.xxz_mpo(n, delta)
        e0 = o.ExactDiagonalization(mpo).evals[0]
        d = o.OracleDMRG2(mpo, o.random_mps(n, chi, 2, seed=n), chi=chi)
        es = [d.sweep(o.RIGHTWARD if k % 2 == 0 else o.LEFTWARD, tol=1e-12) for k in range(4)]
        assert abs(es[-1] - e0) < 1e-10

    def test_dmrg2_truncated_bond_is_variational(self):
        # truncation to chi below the exact bond dims: the energy stays above
        # E0 and the bonds never exceed chi
        n, chi = 10, 4
        mpo = o.random_heisenberg_mpo(n, 0.5, seed=3)
        e0 = o.ExactDiagonalization(mpo).evals[0]
        d = o.OracleDMRG2(mpo, o.random_mps(n, chi, 2, seed=1), chi=chi)
        for k in range(4):
            e = d.sweep(o.RIGHTWARD if k % 2 == 0 else o.LEFTWARD, tol=1e-10)
        assert e >= e0 - 1e-12 and e - e0 < 5e-2
        assert max(a.shape[2] for a in d.mps[:-1]) <= chi

    @pytest.mark.parametrize("direction", [o.RIGHTWARD, o.LEFTWARD])
    def test_split_two_site_eckart_young(self, direction):
        rng = np.random.default_rng(5)
        l, d, r, chi = 6, 2, 5, 4
        theta = rng.standard_normal((l, d, d, r))
        A, B, s = o.split_two_site_svd(theta, chi, direction)
        assert A.shape == (l, d, chi) and B.shape == (chi, d, r) and s.shape == (chi,)
        approx = np.tensordot(A, B, axes=(2, 0))
        # discarded weight: |theta - A B|^2 = |theta|^2 - sum of kept s^2
        resid = np.sum((theta - approx) ** 2)
        assert abs(resid - (np.sum(theta**2) - np.sum(s**2))) < 1e-12
        assert np.all(np.diff(s) <= 0)
        if direction == o.RIGHTWARD:
            Am = A.reshape(l * d, chi)
            np.testing.assert_allclose(Am.T @ Am, np.eye(chi), atol=1e-13)
        else:
            Bm = B.reshape(chi, d * r)
            np.testing.assert_allclose(Bm @ Bm.T, np.eye(chi), atol=1e-13)
        # no other rank-chi factorisation does better (random rank-chi projections)
        M = theta.reshape(l * d, d * r)
        for _ in range(5):
            q, _ = np.linalg.qr(rng.standard_normal((l * d, chi)))
            assert np.sum((M - q @ (q.T @ M)) ** 2) >= resid - 1e-12

    def test_split_two_site_untruncated_is_exact(self):
        theta = np.random.default_rng(6).standard_normal((3, 2, 2, 3))
        A, B, _ = o.split_two_site_svd(theta, 100, o.RIGHTWARD)
        np.testing.assert_allclose(np.tensordot(A, B, axes=(2, 0)), theta, atol=1e-13)


class TestSchmidtOracle:
    """schmidt_values pinned by the dense state's SVD at every bond and by
    the reference's ED formula (tnpy/exact_diagonalization.py:62-90, squared
    singular values of the reshaped eigenvector), itself checked against the
    eigenvalues of the partial trace."""

    def test_matches_dense_svd_every_bond(self):
        n = 8
        mps = o.random_mps(n, 6, 2, seed=21)
        mps[3] = mps[3] * 3.7  # unnormalised, as the DMRG state is
        v = o.mps_to_dense(mps)
        v = v / np.linalg.norm(v)
        sv = o.schmidt_values(mps)
        for i in range(n - 1):
            ref = np.linalg.svd(v.reshape(2 ** (i + 1), -1), compute_uv=False)
            k = len(sv[i])
            np.testing.assert_allclose(sv[i], ref[:k], atol=1e-13)
            assert np.all(ref[k:] < 1e-12)

    def test_unnormalised_scale(self):
        mps = o.random_mps(6, 4, 2, seed=22)
        mps[0] = mps[0] * 2.0
        a = o.schmidt_values(mps, normalize=True)
        b = o.schmidt_values(mps, normalize=False)
        for x, y in zip(a, b):
            np.testing.assert_allclose(2.0 * x, y, atol=1e-13)

    def test_reduced_density_matrix_is_partial_trace(self):
        ed = o.ExactDiagonalization(o.random_heisenberg_mpo(6, 1.0, seed=4))
        psi = ed.evecs[:, 0]
        for site in range(5):
            m = psi.reshape(2 ** (site + 1), -1)
            rho = m @ m.T
            w = np.sort(np.linalg.eigvalsh(rho))[::-1]
            s2 = ed.reduced_density_matrix(site)
            np.testing.assert_allclose(s2, w[: len(s2)], atol=1e-13)
            assert abs(s2.sum() - 1) < 1e-12

    def test_dmrg_ground_state_schmidt_matches_ed(self):
        # a converged one-site DMRG state at exact bond dims carries the ED
        # ground state's Schmidt spectrum (random-field chain: no degeneracy)
        n = 8
        mpo = o.random_heisenberg_mpo(n, 1.0, seed=4)
        ed = o.ExactDiagonalization(mpo)
        d = o.OracleDMRG(mpo, o.random_mps(n, 16, 2, seed=2))
        d.run(tol=1e-13, max_sweep=12, with_variance=False)
        sv = o.schmidt_values(d.mps)
        for i in range(n - 1):
            s2 = ed.reduced_density_matrix(i)
            k = min(len(s2), len(sv[i]))
            np.testing.assert_allclose(sv[i][:k] ** 2, s2[:k], atol=1e-9)


def test_two_site_matvec_equals_einsum_definition():
    """The pairwise-contraction form equals the one-expression definition
    y = sum L W1 W2 R x (small sizes, where einsum is quick)."""
    rng = np.random.default_rng(8)
    L, R = rng.standard_normal((3, 5, 3)), rng.standard_normal((4, 6, 4))
    W1, W2 = rng.standard_normal((5, 7, 2, 2)), rng.standard_normal((7, 6, 2, 2))
    x = rng.standard_normal((3, 2, 2, 4))
    ref = np.einsum("awb,wukc,uvlt,rvs,aklr->bcts", L, W1, W2, R, x).reshape(-1)
    np.testing.assert_allclose(o.two_site_matvec(L, W1, W2, R, x), ref, atol=1e-12)
