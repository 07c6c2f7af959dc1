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
