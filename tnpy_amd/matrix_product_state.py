"""Matrix product state and DMRG environments, resident in HBM.

Mirrors ``tnpy/matrix_product_state.py``:

* ``Direction`` (``:19-25``)
* ``MatrixProductState`` (``:28-228``): site tensors ``'lpr'``; the edge
  sites are stored with a bond of size 1 on the device and exposed in the
  reference's 2-leg shapes through ``mps[site].data``.
* ``Environment`` (``:231-440``): left/right environments ``L[l, w, l']`` /
  ``R[r, w, r']`` (ket, mpo, bra), the one-site matvec, the dense effective
  matrix, the gauge move and the variance, all through libtnx kernels.
* ``MatrixProductStateMeasurements`` (``:443-453``).

The returned state is a quimb-compatible object only when quimb is
installed (``to_quimb``); otherwise this duck-typed MPS carries the same
data, shapes and on-disk format.
"""
from __future__ import annotations

import ctypes as C
from enum import Enum
from pathlib import Path
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from tnpy_amd import _lib
from tnpy_amd.operators import MatrixProductOperator, as_mpo


class Direction(Enum):
    RIGHTWARD = 1
    LEFTWARD = -1


class SiteTensor:
    """View of one site, with the reference's shapes: ``(p, r)`` at site 0,
    ``(l, p)`` at the last site, ``(l, p, r)`` in the bulk."""

    def __init__(self, mps: "MatrixProductState", site: int):
        self._mps = mps
        self._site = site

    @property
    def shape(self):
        l, p, r = self._mps._tensors[self._site].shape
        if self._site == 0:
            return (p, r)
        if self._site == self._mps.n_sites - 1:
            return (l, p)
        return (l, p, r)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))

    @property
    def data(self) -> np.ndarray:
        return self._mps._tensors[self._site].detach().cpu().numpy().reshape(self.shape)

    @property
    def tags(self):
        return {f"I{self._site}"}

    def modify(self, data) -> None:
        self._mps.set_site(self._site, data)


def _bond_dims(n: int, chi: int, d: int) -> List[int]:
    """``min(d^(i+1), d^(n-i-1), chi)`` for bond i, the shape of a compressed
    random MPS (``tests/test_matrix_product_state.py:11-37``)."""
    return [int(min(d ** (i + 1), d ** (n - i - 1), chi)) for i in range(n - 1)]


class MatrixProductState:
    """Finite MPS with site tensors ``A[l, p, r]`` on the GPU.

    Args:
        arrays: per-site arrays (numpy or torch) in ``'lpr'`` order; edge
            sites may be 2-leg (reference shape) or 3-leg with size-1 bond.
        device: torch device (default: the current GPU).
    """

    def __init__(self, arrays: Sequence, device: Optional[torch.device] = None, shape: str = "lpr"):
        if shape != "lpr":
            raise ValueError("only the 'lpr' layout is supported")
        self.device = device if device is not None else _lib.require_gpu()
        n = len(arrays)
        if n < 2:
            raise ValueError("an MPS needs at least 2 sites")
        tensors = []
        for i, a in enumerate(arrays):
            t = torch.as_tensor(np.asarray(a) if not isinstance(a, torch.Tensor) else a)
            t = t.to(device=self.device, dtype=torch.float64)
            if t.dim() == 2 and i == 0:
                t = t.unsqueeze(0)
            elif t.dim() == 2 and i == n - 1:
                t = t.unsqueeze(-1)
            if t.dim() != 3:
                raise ValueError(f"site {i} has shape {tuple(t.shape)}")
            tensors.append(t.contiguous())
        for i in range(n - 1):
            if tensors[i].shape[2] != tensors[i + 1].shape[0]:
                raise ValueError(f"bond mismatch between sites {i} and {i + 1}")
        self._tensors: List[torch.Tensor] = tensors

    # -- reference properties ------------------------------------------------
    @property
    def n_sites(self) -> int:
        return len(self._tensors)

    @property
    def nsites(self) -> int:
        return len(self._tensors)

    @property
    def phys_dim(self) -> int:
        return int(self._tensors[0].shape[1])

    @property
    def bond_dim(self) -> int:
        return max(int(t.shape[2]) for t in self._tensors[:-1])

    def bond_dims(self) -> List[int]:
        return [int(t.shape[2]) for t in self._tensors[:-1]]

    def max_bond(self) -> int:
        return self.bond_dim

    def __getitem__(self, site: int) -> SiteTensor:
        if site < 0:
            site += self.n_sites
        return SiteTensor(self, site)

    def __iter__(self):
        return (SiteTensor(self, i) for i in range(self.n_sites))

    def __len__(self) -> int:
        return self.n_sites

    @property
    def tensors(self) -> List[torch.Tensor]:
        """Device site tensors, always 3-leg ``(l, p, r)``."""
        return self._tensors

    @property
    def arrays(self) -> List[np.ndarray]:
        return [self[i].data for i in range(self.n_sites)]

    def set_site(self, site: int, data) -> None:
        t = torch.as_tensor(data).to(device=self.device, dtype=torch.float64)
        shape = self._tensors[site].shape
        if t.numel() != shape.numel():
            raise ValueError(f"site {site}: expected {tuple(shape)} elements, got {tuple(t.shape)}")
        self._tensors[site] = t.reshape(shape).contiguous()

    def conj(self, mangle_inner: bool = False, mangle_outer: bool = False) -> "MatrixProductState":
        """Real FP64 state: the conjugate is a copy (``:93-116``)."""
        return MatrixProductState([t.clone() for t in self._tensors], device=self.device)

    def copy(self) -> "MatrixProductState":
        return self.conj()

    # -- construction ----------------------------------------------------------
    @classmethod
    def random(
        cls,
        n: int,
        bond_dim: int,
        phys_dim: int,
        seed: Optional[int] = None,
        device: Optional[torch.device] = None,
        host_rng: Optional[bool] = None,
    ) -> "MatrixProductState":
        """Random normalised MPS, right-canonical with the centre at site 0.

        Entries are N(0, 1): drawn on the host from
        ``numpy.random.default_rng(seed)`` site by site in C order of
        ``(l, p, r)`` (``host_rng=True``, the default below 2^26 entries),
        or on the GPU from a seeded ``torch.Generator`` for large states.
        Then an LQ sweep from the right end and a normalisation, the state
        ``MPS_rand_state(...).compress()`` gives at
        ``tnpy/matrix_product_state.py:183-185`` (whose draw is unseeded).
        """
        device = device if device is not None else _lib.require_gpu()
        dims = [1] + _bond_dims(n, bond_dim, phys_dim) + [1]
        shapes = [(dims[i], phys_dim, dims[i + 1]) for i in range(n)]
        total = sum(int(np.prod(s)) for s in shapes)
        if host_rng is None:
            host_rng = total <= (1 << 26)
        if host_rng:
            rng = np.random.default_rng(seed)
            arrays = [rng.standard_normal(s) for s in shapes]
            mps = cls(arrays, device=device)
        else:
            gen = torch.Generator(device=device)
            gen.manual_seed(0 if seed is None else int(seed))
            arrays = [torch.randn(s, generator=gen, device=device, dtype=torch.float64) for s in shapes]
            mps = cls(arrays, device=device)
        mps.right_canonicalize()
        return mps

    def grow(self, bond_dim: int) -> "MatrixProductState":
        """The same state with bonds widened to ``min(d^(i+1), d^(n-i-1),
        bond_dim)`` (the bond-dimension ramp of a DMRG schedule): each site
        tensor is zero-padded into the wider shape, then the state is
        right-canonicalised (Householder LQ completes every padded bond to
        an orthonormal basis) and normalised. Returns a new MPS."""
        d = self.phys_dim
        dims = [1] + _bond_dims(self.n_sites, bond_dim, d) + [1]
        out = []
        for i, t in enumerate(self._tensors):
            l, p, r = t.shape
            L, R = max(dims[i], l), max(dims[i + 1], r)
            big = torch.zeros(L, p, R, dtype=torch.float64, device=self.device)
            big[:l, :, :r] = t
            out.append(big)
        mps = MatrixProductState(out, device=self.device)
        mps.right_canonicalize()
        return mps

    def right_canonicalize(self, normalize: bool = True) -> None:
        """LQ sweep from the right end on the device (tnx_gauge_left)."""
        ws = _lib.Workspace(self.device)
        stream = _lib.stream_ptr(self.device)
        for i in range(self.n_sites - 1, 0, -1):
            self._gauge_left(i, ws, stream)
            if normalize:
                # keep the absorbed factor O(1): a long chain of random
                # tensors otherwise overflows before the final normalisation
                t = self._tensors[i - 1]
                self._tensors[i - 1] = (t / torch.linalg.vector_norm(t)).contiguous()
        if normalize:
            nrm = torch.linalg.vector_norm(self._tensors[0])
            self._tensors[0] = (self._tensors[0] / nrm).contiguous()

    def _gauge_right(self, site: int, ws: "_lib.Workspace", stream: int) -> None:
        A = self._tensors[site]
        l, p, r = A.shape
        nxt = self._tensors[site + 1]
        nl, npp, nr = nxt.shape
        out = torch.empty_like(nxt)
        wp, wb = ws.get(_lib.lib().tnx_gauge_workspace_bytes(l, p, r))
        _lib.check(
            _lib.lib().tnx_gauge_right(
                _lib.ptr(A, "site"), _lib.ptr(nxt, "next"), _lib.ptr(out, "next_out"),
                l, p, r, npp * nr, wp, wb, stream,
            ),
            "tnx_gauge_right",
        )
        self._tensors[site + 1] = out

    def _gauge_left(self, site: int, ws: "_lib.Workspace", stream: int) -> None:
        A = self._tensors[site]
        l, p, r = A.shape
        prv = self._tensors[site - 1]
        pl, pp, pr = prv.shape
        out = torch.empty_like(prv)
        wp, wb = ws.get(_lib.lib().tnx_gauge_workspace_bytes(l, p, r))
        _lib.check(
            _lib.lib().tnx_gauge_left(
                _lib.ptr(A, "site"), _lib.ptr(prv, "prev"), _lib.ptr(out, "prev_out"),
                l, p, r, pl * pp, wp, wb, stream,
            ),
            "tnx_gauge_left",
        )
        self._tensors[site - 1] = out

    def split_tensor(self, site: int, direction: Direction, workspace: Optional["_lib.Workspace"] = None) -> None:
        """Move the orthogonality centre one site (``:187-225``). The
        reference's untruncated SVD becomes a QR (rightward) / LQ (leftward):
        the same state, another gauge of the bond."""
        ws = workspace if workspace is not None else _lib.Workspace(self.device)
        stream = _lib.stream_ptr(self.device)
        if direction == Direction.RIGHTWARD:
            self._gauge_right(site, ws, stream)
        elif direction == Direction.LEFTWARD:
            self._gauge_left(site, ws, stream)
        else:
            raise KeyError("MatrixProductState only supplies left or right direction.")

    # -- measurement helpers ---------------------------------------------------
    def norm(self) -> float:
        """<psi|psi>^(1/2), by a left transfer sweep on the device."""
        E = torch.ones(1, 1, dtype=torch.float64, device=self.device)
        for A in self._tensors:
            l, p, r = A.shape
            # E'[r, r'] = sum A[l,p,r] E[l,l'] A[l',p,r']
            T = (E @ A.reshape(l, p * r)).reshape(l, p, r)
            E = A.reshape(l * p, r).T @ T.reshape(l * p, r)
        return float(E.reshape(-1)[0].clamp_min(0.0).sqrt())

    def schmidt_values(self, normalize: bool = True) -> List[np.ndarray]:
        """Singular values across every bond (gauge invariant), from the
        R factors of a left QR sweep of a right-canonicalised copy (each R
        through the library's SVD). With
        ``normalize`` the state is scaled to unit norm first (the DMRG state
        is not normalised, ``tnpy/finite_dmrg.py:139-141``)."""
        work = self.copy()
        work.right_canonicalize(normalize=normalize)
        out = []
        ws = _lib.Workspace(self.device)
        stream = _lib.stream_ptr(self.device)
        for i in range(self.n_sites - 1):
            A = work._tensors[i]
            l, p, r = A.shape
            nxt = work._tensors[i + 1]
            # QR of the centre; s(R) are the Schmidt values of bond i
            R = torch.empty(r, r, dtype=torch.float64, device=self.device)
            wp, wb = ws.get(_lib.lib().tnx_qr_workspace_bytes(l * p, r))
            _lib.check(_lib.lib().tnx_qr(_lib.ptr(A, "A"), l * p, r, _lib.ptr(R, "R"), wp, wb, stream), "tnx_qr")
            out.append(_lib.svd_trunc(R, r, ws)[2])
            work._tensors[i + 1] = (R @ nxt.reshape(r, -1)).reshape(nxt.shape).contiguous()
        return out

    def to_dense(self) -> np.ndarray:
        """Full state vector (small n only)."""
        v = self._tensors[0].reshape(-1, self._tensors[0].shape[2])
        for A in self._tensors[1:]:
            l, p, r = A.shape
            v = (v @ A.reshape(l, p * r)).reshape(-1, r)
        return v.reshape(-1).cpu().numpy()

    # -- persistence (tnpy/matrix_product_state.py:118-168) ----------------------
    def save(self, filename: str) -> None:
        """Write ``I{site}`` -> ``'lpr'`` arrays (reference shapes) to
        ``.npz`` (or ``.hdf5`` when h5py is installed)."""
        data = {f"I{i}": self[i].data for i in range(self.n_sites)}
        path = Path(filename)
        if path.suffix == ".npz":
            np.savez(str(path), **data)
        elif path.suffix == ".hdf5":
            try:
                import h5py
            except ImportError as e:  # pragma: no cover - h5py is optional
                raise ValueError("File extension .hdf5 needs h5py, which is not installed.") from e
            with h5py.File(str(path), "w") as f:
                for k, v in data.items():
                    f.create_dataset(k, data=v)
        else:
            raise ValueError(f"File extension {path.suffix} is not supported.")

    @classmethod
    def load(cls, filename: str, device: Optional[torch.device] = None) -> "MatrixProductState":
        """Read a file written by ``save``. Sites are ordered by the integer
        in their ``I{site}`` key and read as ``'lpr'`` (the reference's
        loader passes them to quimb's default ``'lrp'`` constructor, the
        ordering issue marked @TODO at ``:130,158``)."""
        path = Path(filename)
        if path.suffix == ".npz":
            with np.load(str(path), allow_pickle=False) as f:
                items = {k: f[k] for k in f.files}
        elif path.suffix == ".hdf5":
            try:
                import h5py
            except ImportError as e:  # pragma: no cover
                raise ValueError("File extension .hdf5 needs h5py, which is not installed.") from e
            with h5py.File(str(path), "r") as f:
                items = {k: f[k][()] for k in f.keys()}
        else:
            raise ValueError(f"File extension {path.suffix} is not supported.")
        keys = sorted(items, key=lambda k: int(k[1:]))
        return cls([items[k] for k in keys], device=device)

    def to_quimb(self):
        """A ``quimb.tensor.MatrixProductState`` with ``'lpr'`` arrays, when
        quimb is installed."""
        try:
            import quimb.tensor as qtn
        except ImportError as e:
            raise ImportError("quimb is not installed; use the arrays of this MPS instead") from e
        return qtn.MatrixProductState(self.arrays, shape="lpr")


class Environment:
    """Left/right environments and one-site effective operators on the GPU
    (``tnpy/matrix_product_state.py:231-440``).

    ``left[i]`` is the environment of site i built from sites 0..i-1
    (``L[0]`` is the trivial 1x1x1 boundary), ``right[i]`` from sites
    i+1..n-1. The constructor builds all right environments; the left ones
    are built as the first rightward sweep reaches them (the reference also
    builds every left environment at construction, ``:247-248``, but each
    is recomputed before first use).
    """

    def __init__(self, mpo, mps: MatrixProductState, build_left: bool = False):
        self._mpo = as_mpo(mpo)
        self._mps = mps
        if self._mpo.n_sites != mps.n_sites:
            raise ValueError("MPO and MPS lengths differ")
        self.device = mps.device
        self._n = mps.n_sites
        self._W = [torch.as_tensor(w, device=self.device).contiguous() for w in self._mpo.padded]
        self._ws = _lib.Workspace(self.device)
        one = torch.ones(1, 1, 1, dtype=torch.float64, device=self.device)
        self._left: Dict[int, torch.Tensor] = {0: one}
        self._right: Dict[int, torch.Tensor] = {self._n - 1: one.clone()}
        self._lib = _lib.lib()
        for site in range(self._n - 2, -1, -1):
            self.update_right(site)
        if build_left:
            for site in range(1, self._n):
                self.update_left(site)

    @property
    def mpo(self) -> MatrixProductOperator:
        return self._mpo

    @property
    def mps(self) -> MatrixProductState:
        return self._mps

    @property
    def n_sites(self) -> int:
        return self._n

    @property
    def left(self) -> Dict[int, torch.Tensor]:
        return self._left

    @property
    def right(self) -> Dict[int, torch.Tensor]:
        return self._right

    @property
    def workspace(self) -> "_lib.Workspace":
        return self._ws

    def close(self) -> None:
        self._left.clear()
        self._right.clear()

    def _stream(self) -> int:
        return _lib.stream_ptr(self.device)

    def update_left(self, site: int, precision: str = "fp64") -> None:
        """``L[site]`` from ``L[site-1]``, A and W of site-1 (``:296-315``)."""
        L = self._left[site - 1]
        A = self._mps.tensors[site - 1]
        W = self._W[site - 1]
        l, d, r = A.shape
        Dl, Dr = W.shape[0], W.shape[1]
        out = torch.empty(r, Dr, r, dtype=torch.float64, device=self.device)
        wp, wb = self._ws.get(self._lib.tnx_env_workspace_bytes(l, Dl, Dr, d, r))
        _lib.check(
            self._lib.tnx_env_left(
                _lib.ptr(L, "L"), _lib.ptr(A, "A"), _lib.ptr(W, "W"), _lib.ptr(out, "Lout"),
                l, Dl, Dr, d, r, _lib.PRECISIONS[precision], wp, wb, self._stream(),
            ),
            "tnx_env_left",
        )
        self._left[site] = out

    def update_right(self, site: int, precision: str = "fp64") -> None:
        """``R[site]`` from ``R[site+1]``, A and W of site+1 (``:317-336``)."""
        R = self._right[site + 1]
        A = self._mps.tensors[site + 1]
        W = self._W[site + 1]
        l, d, r = A.shape
        Dl, Dr = W.shape[0], W.shape[1]
        out = torch.empty(l, Dl, l, dtype=torch.float64, device=self.device)
        wp, wb = self._ws.get(self._lib.tnx_env_workspace_bytes(l, Dl, Dr, d, r))
        _lib.check(
            self._lib.tnx_env_right(
                _lib.ptr(R, "R"), _lib.ptr(A, "A"), _lib.ptr(W, "W"), _lib.ptr(out, "Rout"),
                l, Dl, Dr, d, r, _lib.PRECISIONS[precision], wp, wb, self._stream(),
            ),
            "tnx_env_right",
        )
        self._right[site] = out

    def update(self, site: int, direction: Direction) -> None:
        if direction == Direction.RIGHTWARD:
            self.update_left(site + 1)
        elif direction == Direction.LEFTWARD:
            self.update_right(site - 1)

    def ensure_left(self, site: int) -> None:
        """Build missing left environments up to ``site``."""
        first = max(k for k in self._left if k <= site)
        for s in range(first + 1, site + 1):
            self.update_left(s)

    def update_mps(self, site: int, data) -> None:
        self._mps.set_site(site, data)

    def split_tensor(self, site: int, direction: Direction) -> None:
        self._mps.split_tensor(site, direction, workspace=self._ws)

    def site_dims(self, site: int):
        A = self._mps.tensors[site]
        W = self._W[site]
        l, d, r = A.shape
        return l, int(W.shape[0]), int(W.shape[1]), d, r

    def matvec_device(self, site: int, x: torch.Tensor, y: torch.Tensor, precision: str = "fp64") -> None:
        """y = H_eff(site) x on the device (``:411-440``)."""
        l, Dl, Dr, d, r = self.site_dims(site)
        wp, wb = self._ws.get(self._lib.tnx_heff1_workspace_bytes(l, Dl, Dr, d, r))
        _lib.check(
            self._lib.tnx_heff1_matvec(
                _lib.ptr(self._left[site], "L"), _lib.ptr(self._W[site], "W"),
                _lib.ptr(self._right[site], "R"), _lib.ptr(x, "x"), _lib.ptr(y, "y"),
                l, Dl, Dr, d, r, _lib.PRECISIONS[precision], wp, wb, self._stream(),
            ),
            "tnx_heff1_matvec",
        )

    def one_site_matvec(self, site: int):
        """A scipy ``LinearOperator`` over host vectors, the reference's
        operator seam (``:411-440``); each call round-trips to the GPU."""
        from scipy.sparse.linalg import LinearOperator

        N = int(self._mps.tensors[site].numel())

        def mv(x):
            xt = torch.as_tensor(np.ascontiguousarray(np.ravel(x)), device=self.device, dtype=torch.float64)
            yt = torch.empty_like(xt)
            self.matvec_device(site, xt, yt)
            return yt.cpu().numpy().reshape(-1, 1)

        return LinearOperator(shape=(N, N), matvec=mv, dtype=float)

    def full_matrix_device(self, site: int) -> torch.Tensor:
        l, Dl, Dr, d, r = self.site_dims(site)
        N = l * d * r
        M = torch.empty(N, N, dtype=torch.float64, device=self.device)
        _lib.check(
            self._lib.tnx_heff1_dense(
                _lib.ptr(self._left[site], "L"), _lib.ptr(self._W[site], "W"),
                _lib.ptr(self._right[site], "R"), _lib.ptr(M, "M"), l, Dl, Dr, d, r, self._stream(),
            ),
            "tnx_heff1_dense",
        )
        return M

    def one_site_full_matrix(self, site: int) -> np.ndarray:
        """Dense H_eff (``:372-409``), rows (l,k,r) ket, columns (l',b,r') bra."""
        return self.full_matrix_device(site).cpu().numpy()

    def expectation(self, mpo: Optional[MatrixProductOperator] = None) -> float:
        """<psi|O|psi> (unnormalised) by a left sweep of env updates with O's
        tensors (O = this environment's MPO by default)."""
        mpo = self._mpo if mpo is None else as_mpo(mpo)
        Ws = [torch.as_tensor(w, device=self.device).contiguous() for w in mpo.padded]
        L = torch.ones(1, 1, 1, dtype=torch.float64, device=self.device)
        ws = self._ws
        for site, (A, W) in enumerate(zip(self._mps.tensors, Ws)):
            l, d, r = A.shape
            Dl, Dr = int(W.shape[0]), int(W.shape[1])
            out = torch.empty(r, Dr, r, dtype=torch.float64, device=self.device)
            wp, wb = ws.get(self._lib.tnx_env_workspace_bytes(l, Dl, Dr, d, r))
            _lib.check(
                self._lib.tnx_env_left(
                    _lib.ptr(L, "L"), _lib.ptr(A, "A"), _lib.ptr(W, "W"), _lib.ptr(out, "Lout"),
                    l, Dl, Dr, d, r, _lib.TNX_F64, wp, wb, self._stream(),
                ),
                "tnx_env_left",
            )
            L = out
        return float(L.reshape(-1)[0])

    def variance(self) -> float:
        """<psi|H^2|psi> - <psi|H|psi>^2, unnormalised (``:367-370``), with
        the squared MPO of ``tnpy/operators.py:89-116``."""
        if not hasattr(self, "_mpo_sq"):
            self._mpo_sq = self._mpo.square()
        return self.expectation(self._mpo_sq) - self.expectation() ** 2


class MatrixProductStateMeasurements:
    """``tnpy/matrix_product_state.py:443-453``: unnormalised expectation
    values of an MPO (or the norm squared without one)."""

    def __init__(self, mps: MatrixProductState):
        self._mps = mps

    def expectation_value(self, mpo=None) -> float:
        if mpo is None:
            return self._mps.norm() ** 2
        env = Environment.__new__(Environment)
        env._mps = self._mps
        env.device = self._mps.device
        env._ws = _lib.Workspace(self._mps.device)
        env._lib = _lib.lib()
        env._mpo = as_mpo(mpo)
        return Environment.expectation(env, mpo)
