"""Host-side parts of tnpy_amd that run without a GPU: the model MPOs
against the reference's literal Hamiltonians, the MPO square, and the
C-ABI library's exported symbols and argument validation."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import dmrg_ref as o
from tnpy_amd.model import XXZ, RandomCouplingHeisenberg, RandomHeisenberg
from tnpy_amd.operators import MatrixProductOperator, SpinOperators

ROOT = Path(__file__).resolve().parents[1]


def dense(mpo: MatrixProductOperator) -> np.ndarray:
    return o.full_hamiltonian(mpo.padded)


class TestModels:
    @pytest.mark.parametrize(
        "key, model",
        [
            ("random_heisenberg_n2_h0", lambda: RandomHeisenberg(n=2, h=0)),
            ("random_heisenberg_n3_h0", lambda: RandomHeisenberg(n=3, h=0)),
            ("xxz_n2_delta0.5", lambda: XXZ(n=2, delta=0.5)),
            ("xxz_n3_delta0.5", lambda: XXZ(n=3, delta=0.5)),
        ],
    )
    def test_literal_hamiltonians(self, fixtures, key, model):
        np.testing.assert_array_equal(dense(model().mpo), np.array(fixtures["operators"][key]))

    def test_reference_edge_shapes(self):
        mpo = RandomHeisenberg(n=4, h=0.5, seed=1).mpo
        assert mpo[0].shape == (5, 2, 2) and mpo[1].shape == (5, 5, 2, 2) and mpo[3].shape == (5, 2, 2)
        assert mpo.n_sites == 4 and mpo.phys_dim == 2

    def test_square(self):
        mpo = RandomHeisenberg(n=4, h=0.5, seed=2).mpo
        sq = mpo.square()
        assert sq[0].shape == (25, 2, 2) and sq[1].shape == (25, 25, 2, 2)
        H = dense(mpo)
        np.testing.assert_allclose(H @ H, dense(sq), atol=1e-12)

    @pytest.mark.parametrize("n", [2, 4, 6])
    def test_multiply_scalar(self, n):
        mpo = RandomHeisenberg(n=n, h=0.5, seed=3).mpo
        np.testing.assert_array_equal(-1 * dense(mpo), dense(-1 * mpo))

    def test_notebook_spectrum(self, fixtures):
        H = dense(RandomHeisenberg(n=6, h=3.0, seed=2021).mpo)
        np.testing.assert_allclose(
            np.linalg.eigvalsh(H), fixtures["notebook_random_heisenberg_n6_h3_seed2021_evals"], atol=6e-8
        )

    def test_random_coupling_matches_oracle(self):
        m = RandomCouplingHeisenberg(n=7, seed=11)
        np.testing.assert_array_equal(
            dense(m.mpo), o.full_hamiltonian(o.random_coupling_heisenberg_mpo(7, m.couplings))
        )

    def test_spin_operators(self):
        Sp, Sm, Sz, I2, O2 = SpinOperators()
        np.testing.assert_array_equal(Sp + Sm, [[0, 1], [1, 0]])

    def test_mpo_validation(self):
        with pytest.raises(ValueError):
            MatrixProductOperator([np.zeros((5, 2, 2)), np.zeros((4, 2, 2))])


def _declared_symbols():
    text = (ROOT / "include" / "tnx.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tnx_[a-z0-9_]+)\s*\(", text)))


class TestABI:
    @pytest.fixture(scope="class")
    def lib(self):
        from tnpy_amd import _lib

        return _lib.load_library()

    def test_exports_every_declared_symbol(self, lib):
        syms = _declared_symbols()
        assert len(syms) >= 20
        for s in syms:
            assert hasattr(lib, s), f"{s} declared in include/tnx.h but not exported"

    def test_python_binding_covers_header(self):
        from tnpy_amd import _lib

        assert set(_declared_symbols()) == set(_lib.SIGNATURES)

    def test_workspace_queries(self, lib):
        assert lib.tnx_heff1_workspace_bytes(1024, 5, 5, 2, 1024) >= (2 * 1024 * 10 * 1024) * 8
        assert lib.tnx_qr_workspace_bytes(2048, 1024) >= 2048 * 1024 * 8
        assert lib.tnx_syev_lowest_workspace_bytes(64) > 0

    def test_invalid_arguments_return_status(self, lib):
        # validated on the host before any device work
        st = lib.tnx_heff1_matvec(None, None, None, None, None, 4, 5, 5, 2, 4, 0, None, 0, None)
        assert st == -1
        assert b"null" in lib.tnx_last_error()
        st = lib.tnx_heff1_matvec(None, None, None, None, None, 0, 5, 5, 2, 4, 0, None, 0, None)
        assert st == -1 and b"dims" in lib.tnx_last_error()
        assert lib.tnx_qr(None, 2, 4, None, None, 0, None) == -1
        assert lib.tnx_gauge_right(None, None, None, 4, 2, 4, 8, None, 0, None) == -1
        assert lib.tnx_syev_lowest(None, 300, None, None, None, 0, None) == -1
        assert lib.tnx_axpy(-1, 1.0, None, None, None) == -1
        from tnpy_amd import _lib

        opts = _lib.EigshOpts(tol=1e-8, max_basis=32, max_matvecs=100, precision=7)
        info = _lib.EigshInfo()
        st = lib.tnx_eigsh_heff1(1, 1, 1, 1, 1, None, 4, 5, 5, 2, 4, C.byref(opts), C.byref(info), None, 0, None)
        assert st == -1 and b"precision" in lib.tnx_last_error()

    def test_tile_table_round_trip(self, lib, tmp_path):
        """The autotuner's tile table loads valid lines only and saves what it
        holds (host-only: no GEMM runs)."""
        src = tmp_path / "tiles.txt"
        src.write_text(
            "# comment\n"
            "0 2048 5120 1024 1 1 1 0 0 0 1 4\n"
            "1 1024 2048 5120 1 0 1 0 1 0 4 2\n"
            "0 8 8 8 1 1 1 0 0 0 1 99\n"  # no such configuration: skipped
            "garbage\n"
        )
        assert lib.tnx_tune_load(str(src).encode()) == 2
        out = tmp_path / "saved.txt"
        assert lib.tnx_tune_save(str(out).encode()) >= 2
        lines = [l for l in out.read_text().splitlines() if not l.startswith("#")]
        assert "0 2048 5120 1024 1 1 1 0 0 0 1 4" in lines
        assert "1 1024 2048 5120 1 0 1 0 1 0 4 2" in lines
        assert lib.tnx_tune_load(str(tmp_path / "missing.txt").encode()) == -1
        lib.tnx_tune_clear()
        assert lib.tnx_tune_save(str(out).encode()) == 0
        from tnpy_amd import _lib

        assert lib.tnx_tune_load(str(_lib.TILE_TABLE).encode()) > 0  # restore the committed table

    def test_product_raises_without_gpu(self):
        import torch

        if torch.cuda.is_available():
            pytest.skip("GPU present")
        from tnpy_amd import FiniteDMRG

        with pytest.raises(RuntimeError, match="GPU"):
            FiniteDMRG(XXZ(n=6, delta=0.5).mpo, bond_dim=4)


def test_batched_stack_joins_adjacent_views_without_copy():
    """The grouped batched sweep joins the stacks of several length groups;
    stacks that lie back to back in one buffer (the previous step's output)
    are joined as a view, anything else by a copy in the given order."""
    import torch

    from tnpy_amd.batched_dmrg import _stack

    buf = torch.arange(5 * 2 * 3, dtype=torch.float64).reshape(5, 2, 3)
    a, b = buf[0:2], buf[2:5]
    v = _stack([a, b])
    assert v.data_ptr() == buf.data_ptr() and torch.equal(v, buf)
    other = torch.zeros(1, 2, 3, dtype=torch.float64)
    w = _stack([a, other])
    assert w.data_ptr() != buf.data_ptr() and torch.equal(w, torch.cat([a, other]))
    x = _stack([b, a])  # out of order: copied
    assert x.data_ptr() != buf.data_ptr() and torch.equal(x, torch.cat([b, a]))
    assert _stack([a]) is a
