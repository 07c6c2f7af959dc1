"""GPU ExactDiagonalization (tnpy/exact_diagonalization.py) and the dense
symmetric eigensolver behind it, against the reference's literals
(tests/test_exact_diagonalization.py:15-42), the notebook spectrum, the
oracle's dense Hamiltonian and numpy.linalg.eigh."""
import numpy as np
import pytest
import torch

from oracle import dmrg_ref as o

pytestmark = pytest.mark.gpu


def _check_decomposition(H, evals, V, tol):
    """Eigenpairs of the symmetric H: ascending, orthonormal, residuals."""
    nrm = np.abs(H).sum(axis=1).max()
    assert np.all(np.diff(evals) >= -1e-12 * nrm)
    assert np.abs(V.T @ V - np.eye(len(evals))).max() < 1e-12
    assert np.abs(H @ V - V * evals).max() <= tol * nrm


def test_heisenberg_n2_literals(gpu, fixtures):
    # tests/test_exact_diagonalization.py:15-42 (the degenerate triplet's
    # basis is LAPACK's choice; the singlet and the spectrum are fixed)
    from tnpy_amd import ExactDiagonalization
    from tnpy_amd.model import RandomHeisenberg

    ed = ExactDiagonalization(RandomHeisenberg(n=2, h=0).mpo)
    np.testing.assert_array_equal(ed.matrix, np.array(fixtures["ed"]["heisenberg_n2_matrix"]))
    np.testing.assert_allclose(ed.evals, [-0.75, 0.25, 0.25, 0.25], atol=1e-14)
    g = ed.evecs[:, 0]
    assert abs(abs(g @ np.array([0, 1, -1, 0]) / np.sqrt(2)) - 1) < 1e-14
    # the triplet eigenspace: projector equal to the reference evecs' span
    ref = np.array(fixtures["ed"]["heisenberg_n2_evecs"])[:, 1:]
    np.testing.assert_allclose(ed.evecs[:, 1:] @ ed.evecs[:, 1:].T, ref @ ref.T, atol=1e-14)


def test_notebook_spectrum(gpu, fixtures):
    # docs/source/notebooks/tsdrg.ipynb:140-150, printed to 8 digits
    from tnpy_amd import ExactDiagonalization
    from tnpy_amd.model import RandomHeisenberg

    ed = ExactDiagonalization(RandomHeisenberg(n=6, h=3.0, seed=2021).mpo)
    np.testing.assert_allclose(ed.evals, fixtures["notebook_random_heisenberg_n6_h3_seed2021_evals"], atol=6e-8)
    H = o.full_hamiltonian(RandomHeisenberg(n=6, h=3.0, seed=2021).mpo.padded)
    np.testing.assert_array_equal(ed.matrix, H)
    np.testing.assert_allclose(ed.evals, np.linalg.eigvalsh(H), atol=1e-13)


@pytest.mark.parametrize("n, delta", [(8, 0.5), (10, 1.0), (12, 0.5)])
def test_xxz_full_spectrum(gpu, n, delta):
    """Dense Hamiltonian equal to the oracle's; every eigenpair: ascending,
    orthonormal, residual within 1e-11 of |H| (measured 1e-12 at 4096);
    eigenvalues equal numpy's (delta=1
    is the SU(2) ferromagnet: multiplets up to (n+1)-fold degenerate)."""
    import time

    from tnpy_amd import ExactDiagonalization
    from tnpy_amd.model import XXZ

    print(f"ED XXZ n={n} delta={delta}", flush=True)
    t = time.perf_counter()
    ed = ExactDiagonalization(XXZ(n=n, delta=delta).mpo)
    print(f"GPU ED {2**n}x{2**n}: {time.perf_counter() - t:.2f}s", flush=True)
    H = o.full_hamiltonian(o.xxz_mpo(n, delta))
    np.testing.assert_array_equal(ed.matrix, H)
    _check_decomposition(H, ed.evals, ed.evecs, 1e-11)
    np.testing.assert_allclose(ed.evals, np.linalg.eigvalsh(H), atol=1e-12 * np.abs(H).sum(axis=1).max())
    if delta == 1.0:
        assert abs(ed.evals[0] + (n - 1) / 4) < 1e-12


def test_schmidt_and_measurements(gpu):
    """reduced_density_matrix against the oracle's SVD of the same ground
    state and the partial trace; entropy; one-/two-point functions and the
    variance against dense numpy on the returned eigenvectors."""
    from tnpy_amd import ExactDiagonalization
    from tnpy_amd.model import RandomHeisenberg

    model = RandomHeisenberg(n=8, h=0.5, seed=2021)
    ed = ExactDiagonalization(model.mpo)
    ref = o.ExactDiagonalization(model.mpo.padded)
    assert abs(ed.evals[0] - ref.evals[0]) < 1e-13
    psi = ed.evecs[:, 0]
    for site in range(7):
        s2 = ed.reduced_density_matrix(site)
        np.testing.assert_allclose(s2, ref.reduced_density_matrix(site), atol=1e-13)
        m = psi.reshape(2 ** (site + 1), -1)
        w = np.sort(np.linalg.eigvalsh(m @ m.T))[::-1]
        np.testing.assert_allclose(s2, w[: len(s2)], atol=1e-13)
        assert abs(s2.sum() - 1) < 1e-12
    ent = ed.entanglement_entropy(3, nan_to_num=True)
    s2 = ed.reduced_density_matrix(3)
    assert abs(ent - np.nan_to_num(-(s2 @ np.log(s2)))) < 1e-12
    Sp, Sm, Sz, I2, _ = o.spin_operators()
    Oz = np.kron(np.kron(np.eye(2**2), Sz), np.eye(2**5))
    assert abs(ed.one_point_function(Sz, 2) - psi @ Oz @ psi) < 1e-13
    Oc = np.kron(np.kron(np.kron(np.eye(2), Sz), np.eye(2**3)), np.kron(Sz, np.eye(2**2)))
    assert abs(ed.two_point_function(Sz, Sz, 1, 5, 0) - psi @ Oc @ psi) < 1e-13
    conn = ed.connected_two_point_function(Sz, Sz, 1, 5, 0)
    assert abs(conn - (psi @ Oc @ psi - ed.one_point_function(Sz, 1) * ed.one_point_function(Sz, 5))) < 1e-13
    var = ed.variance()
    assert np.abs(var).max() < 1e-10  # eigenstates of H have zero variance
    with pytest.raises(ValueError):
        ed.reduced_density_matrix(7)


def test_projected_and_resource_limit(gpu):
    from tnpy_amd import ExactDiagonalization
    from tnpy_amd.model import XXZ

    n = 6
    H = o.full_hamiltonian(o.xxz_mpo(n, 0.5))
    P, _ = np.linalg.qr(np.random.default_rng(0).standard_normal((2**n, 20)))
    ed = ExactDiagonalization(XXZ(n=n, delta=0.5).mpo, proj=P)
    np.testing.assert_allclose(ed.matrix, P.T @ H @ P, atol=1e-13)
    np.testing.assert_allclose(ed.evals, np.linalg.eigvalsh(P.T @ H @ P), atol=1e-13)
    with pytest.raises(ResourceWarning):
        ExactDiagonalization(XXZ(n=13, delta=0.5).mpo)
