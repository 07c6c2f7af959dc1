"""One-site finite DMRG of many independent chains at once.

The reference sweeps one chain per ``FiniteDMRG`` object
(``tnpy/finite_dmrg.py:24-41``, ``sweep`` at ``:143-171``, ``run`` at
``:214-257``); a study of many disorder realisations is a Python loop over
such objects. ``BatchedFiniteDMRG`` runs the same algorithm for a list of
chains: chains of equal length share every tensor shape at every site, so
each length group keeps its tensors as stacks (chain index first) and runs
each site step for the whole group with one launch per kernel:

- the lowest eigenpair of every chain's H_eff (batched dense eigensolver
  below ``exact_solver_dim``, lockstep batched Lanczos otherwise, where each
  chain converges on its own residual),
- psi <- psi + 1e-5 H psi,
- the QR / LQ gauge move (one workgroup per chain, LDS resident),
- the environment update.

Groups of different lengths step in lockstep (rightward sweeps aligned at
the left end, leftward sweeps at the right end), and the groups whose
current sites have equal shapes are joined into one stack, so a step is one
launch per kernel for all of them; away from the right ends every group has
the same interior shape. Per chain the result equals a ``FiniteDMRG`` run from the same initial
state up to rounding. FP64 throughout.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
import weakref
from concurrent.futures import ThreadPoolExecutor
from itertools import cycle
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from tnpy_amd import _lib, logger
from tnpy_amd.matrix_product_state import Direction, MatrixProductState, _bond_dims
from tnpy_amd.operators import as_mpo


def _env_left(ws, stream, L, A, W, dims, B) -> torch.Tensor:
    """Stacked left environments of the next site (tnx_env_left_batched)."""
    l, Dl, Dr, d, r = dims
    out = torch.empty(B, r, Dr, r, dtype=torch.float64, device=A.device)
    lib = _lib.lib()
    wp, wb = ws.get(lib.tnx_env_batched_workspace_bytes(l, Dl, Dr, d, r, B))
    _lib.check(
        lib.tnx_env_left_batched(
            _lib.ptr(L, "L"), _lib.ptr(A, "A"), _lib.ptr(W, "W"), _lib.ptr(out, "Lout"),
            l, Dl, Dr, d, r, B, _lib.TNX_F64, wp, wb, stream,
        ),
        "tnx_env_left_batched",
    )
    return out


def _env_right(ws, stream, R, A, W, dims, B) -> torch.Tensor:
    """Stacked right environments of the previous site (tnx_env_right_batched)."""
    l, Dl, Dr, d, r = dims
    out = torch.empty(B, l, Dl, l, dtype=torch.float64, device=A.device)
    lib = _lib.lib()
    wp, wb = ws.get(lib.tnx_env_batched_workspace_bytes(l, Dl, Dr, d, r, B))
    _lib.check(
        lib.tnx_env_right_batched(
            _lib.ptr(R, "R"), _lib.ptr(A, "A"), _lib.ptr(W, "W"), _lib.ptr(out, "Rout"),
            l, Dl, Dr, d, r, B, _lib.TNX_F64, wp, wb, stream,
        ),
        "tnx_env_right_batched",
    )
    return out


def _matvec(ws, stream, L, W, R, x, y, dims, B) -> None:
    l, Dl, Dr, d, r = dims
    lib = _lib.lib()
    wp, wb = ws.get(lib.tnx_heff1_batched_workspace_bytes(l, Dl, Dr, d, r, B))
    _lib.check(
        lib.tnx_heff1_matvec_batched(
            _lib.ptr(L, "L"), _lib.ptr(W, "W"), _lib.ptr(R, "R"), _lib.ptr(x, "x"), _lib.ptr(y, "y"),
            l, Dl, Dr, d, r, B, _lib.TNX_F64, wp, wb, stream,
        ),
        "tnx_heff1_matvec_batched",
    )


def _solve(owner, ws, stream, L, W, R, v0, dims, B, chains, site, tol):
    """Lowest eigenpair of each of B stacked H_eff: (energies, vec, hvec),
    vec/hvec stacked (B, N). ``chains`` names the problems in warnings."""
    lib = _lib.lib()
    l, Dl, Dr, d, r = dims
    N = l * d * r
    dev = owner.device
    vec = torch.empty(B, N, dtype=torch.float64, device=dev)
    hvec = torch.empty(B, N, dtype=torch.float64, device=dev)
    evals = (C.c_double * B)()
    if N < owner._exact_solver_dim:
        M = torch.empty(B, N, N, dtype=torch.float64, device=dev)
        _lib.check(
            lib.tnx_heff1_dense_batched(
                _lib.ptr(L, "L"), _lib.ptr(W, "W"), _lib.ptr(R, "R"), _lib.ptr(M, "M"),
                l, Dl, Dr, d, r, B, stream,
            ),
            "tnx_heff1_dense_batched",
        )
        wp, wb = ws.get(lib.tnx_syev_lowest_batched_workspace_bytes(N, B))
        _lib.check(
            lib.tnx_syev_lowest_batched(_lib.ptr(M, "M"), N, B, _lib.ptr(vec, "evec"), evals, wp, wb, stream),
            "tnx_syev_lowest_batched",
        )
        _matvec(ws, stream, L, W, R, vec, hvec, dims, B)
        with owner._lock:
            owner.stats["dense_sites"] += 1
        return np.array(evals[:]), vec, hvec
    opts = _lib.EigshOpts()
    opts.tol = float(tol)
    opts.max_basis = int(owner._max_basis)
    opts.max_matvecs = int(owner._max_matvecs)
    opts.precision = _lib.TNX_F64
    opts.keep = int(owner._keep)
    opts.steps_per_check = int(owner._steps_per_check)
    infos = (_lib.EigshInfo * B)()
    wp, wb = ws.get(lib.tnx_eigsh_heff1_batched_workspace_bytes(l, Dl, Dr, d, r, opts.max_basis, B))
    v0 = v0.reshape(B, N)
    st = lib.tnx_eigsh_heff1_batched(
        _lib.ptr(L, "L"), _lib.ptr(W, "W"), _lib.ptr(R, "R"), _lib.ptr(v0, "v0"), _lib.ptr(vec, "vec"),
        _lib.ptr(hvec, "hvec"), l, Dl, Dr, d, r, B, C.byref(opts), infos, wp, wb, stream,
    )
    if st not in (_lib.TNX_OK, _lib.TNX_ERR_NOCONV):
        _lib.check(st, "tnx_eigsh_heff1_batched")
    if st == _lib.TNX_ERR_NOCONV:
        bad = [chains[i] for i in range(B) if not infos[i].converged]
        logger.warning(f"Lanczos at site {site} hit the matvec limit for chains {bad}.")
    with owner._lock:
        owner.stats["krylov_sites"] += 1
        owner.stats["matvecs"] += sum(infos[i].matvecs for i in range(B))
        owner.stats["lockstep_matvecs"] += max(infos[i].matvecs for i in range(B))
        owner.stats["restarts"] += max(infos[i].restarts for i in range(B))
    return np.array([infos[i].eval for i in range(B)]), vec, hvec


def _gauge_right(ws, stream, A, nxt, B) -> torch.Tensor:
    """QR of each stacked site tensor A (in place) with R absorbed into the
    next site; returns the new next-site stack."""
    lib = _lib.lib()
    _, l, p, r = A.shape
    out = torch.empty_like(nxt)
    wp, wb = ws.get(lib.tnx_gauge_batched_workspace_bytes(l, p, r, B))
    _lib.check(
        lib.tnx_gauge_right_batched(
            _lib.ptr(A, "site"), _lib.ptr(nxt, "next"), _lib.ptr(out, "next_out"),
            l, p, r, nxt.shape[2] * nxt.shape[3], B, wp, wb, stream,
        ),
        "tnx_gauge_right_batched",
    )
    return out


def _gauge_left(ws, stream, A, prv, B) -> torch.Tensor:
    """LQ of each stacked site tensor A (in place) with L absorbed into the
    previous site; returns the new previous-site stack."""
    lib = _lib.lib()
    _, l, p, r = A.shape
    out = torch.empty_like(prv)
    wp, wb = ws.get(lib.tnx_gauge_batched_workspace_bytes(l, p, r, B))
    _lib.check(
        lib.tnx_gauge_left_batched(
            _lib.ptr(A, "site"), _lib.ptr(prv, "prev"), _lib.ptr(out, "prev_out"),
            l, p, r, prv.shape[1] * prv.shape[2], B, wp, wb, stream,
        ),
        "tnx_gauge_left_batched",
    )
    return out


def _stack(parts: List[torch.Tensor]) -> torch.Tensor:
    """The stacks of several groups as one stack (chain index first). When
    the parts already lie back to back in one buffer (the outputs of the
    previous merged step) this is a view of that buffer, else a copy."""
    if len(parts) == 1:
        return parts[0]
    first = parts[0]
    rest = first.shape[1:]
    ok = all(p.is_contiguous() and p.shape[1:] == rest for p in parts)
    if ok:
        st = first.untyped_storage()
        ptr = first.data_ptr()
        for p in parts:
            if p.untyped_storage().data_ptr() != st.data_ptr() or p.data_ptr() != ptr:
                ok = False
                break
            ptr += p.numel() * p.element_size()
    if ok:
        total = sum(p.shape[0] for p in parts)
        return first.new_empty(0).set_(st, first.storage_offset(), (total,) + tuple(rest),
                                       first.stride())
    return torch.cat(parts)


class _Group:
    """Stacked state of the chains of one length."""

    def __init__(self, owner: "BatchedFiniteDMRG", chains: List[int], mpos, stacks: List[torch.Tensor]):
        self.owner = owner
        self.chains = chains
        self.B = len(chains)
        dev = owner.device
        # each length group runs on its own stream (default priority) with
        # its own scratch, so groups can be swept concurrently
        self.stream = torch.cuda.Stream(device=dev)
        self.stream.wait_stream(torch.cuda.current_stream(dev))
        self.ws = _lib.Workspace(dev)
        with torch.cuda.stream(self.stream):
            self._build(mpos, stacks)

    def _build(self, mpos, stacks: List[torch.Tensor]) -> None:
        dev = self.owner.device
        self.n = mpos[0].n_sites
        self.W = [
            torch.stack([torch.as_tensor(m.padded[s], device=dev) for m in mpos]).contiguous()
            for s in range(self.n)
        ]
        self.A = stacks
        one = torch.ones(self.B, 1, 1, 1, dtype=torch.float64, device=dev)
        self.left: Dict[int, torch.Tensor] = {0: one}
        self.right: Dict[int, torch.Tensor] = {self.n - 1: one.clone()}
        for site in range(self.n - 2, -1, -1):
            self.update_right(site)

    # -- building blocks -------------------------------------------------------
    def dims(self, site: int):
        _, l, d, r = self.A[site].shape
        W = self.W[site]
        return l, int(W.shape[1]), int(W.shape[2]), d, r

    def update_left(self, site: int) -> None:
        self.left[site] = _env_left(self.ws, self.stream.cuda_stream, self.left[site - 1], self.A[site - 1],
                                    self.W[site - 1], self.dims(site - 1), self.B)

    def update_right(self, site: int) -> None:
        self.right[site] = _env_right(self.ws, self.stream.cuda_stream, self.right[site + 1], self.A[site + 1],
                                      self.W[site + 1], self.dims(site + 1), self.B)

    def ensure_left(self, site: int) -> None:
        first = max(k for k in self.left if k <= site)
        for s in range(first + 1, site + 1):
            self.update_left(s)

    def matvec(self, site: int, x: torch.Tensor, y: torch.Tensor) -> None:
        _matvec(self.ws, self.stream.cuda_stream, self.left[site], self.W[site], self.right[site], x, y,
                self.dims(site), self.B)

    def solve(self, site: int, tol: float):
        """Lowest eigenpair of every chain's H_eff at ``site``: (energies,
        vec, hvec), vec/hvec stacked (B, N)."""
        return _solve(self.owner, self.ws, self.stream.cuda_stream, self.left[site], self.W[site],
                      self.right[site], self.A[site], self.dims(site), self.B, self.chains, site, tol)

    def gauge(self, site: int, direction: Direction) -> None:
        if direction == Direction.RIGHTWARD:
            self.A[site + 1] = _gauge_right(self.ws, self.stream.cuda_stream, self.A[site], self.A[site + 1], self.B)
        else:
            self.A[site - 1] = _gauge_left(self.ws, self.stream.cuda_stream, self.A[site], self.A[site - 1], self.B)

    # -- site step and sweep -------------------------------------------------------
    def site_step(self, site: int, direction: Direction, tol: float) -> np.ndarray:
        if site > 0 and site not in self.left:
            self.ensure_left(site)
        energies, vec, hvec = self.solve(site, tol)
        _lib.check(
            _lib.lib().tnx_axpy(vec.numel(), 1e-5, _lib.ptr(hvec, "hvec"), _lib.ptr(vec, "vec"),
                                self.stream.cuda_stream),
            "tnx_axpy",
        )
        self.A[site] = vec.reshape(self.A[site].shape)
        self.gauge(site, direction)
        if direction == Direction.RIGHTWARD:
            self.update_left(site + 1)
        else:
            self.update_right(site - 1)
        return energies

    def sweep(self, direction: Direction, tol: float) -> np.ndarray:
        # worker threads start on the default device; make it this group's
        torch.cuda.set_device(self.owner.device)
        with torch.cuda.stream(self.stream):
            if direction == Direction.LEFTWARD:
                self.ensure_left(self.n - 1)
            sites = range(self.n - 1) if direction == Direction.RIGHTWARD else range(self.n - 1, 0, -1)
            energies = None
            for site in sites:
                energies = self.site_step(site, direction, tol)
            self.stream.synchronize()
            return energies


class BatchedFiniteDMRG:
    """One-site finite DMRG of a list of independent chains.

    Args:
        mpos: one MPO per chain (``MatrixProductOperator``, W arrays or a
            quimb MPO). Lengths may differ; the physical dimension may not.
        bond_dim: bond dimension of every chain (``chi`` is an alias).
        mps: optional initial states, one per chain (``MatrixProductState``
            or arrays, 'lpr' site tensors).
        seeds: seeds of the random initial states (default 0, 1, 2, ...),
            each drawn as ``MatrixProductState.random(n, chi, d, seed)``.
        exact_solver_dim, max_basis, keep: as for ``FiniteDMRG``.
        max_matvecs: Lanczos matvec limit per site.
        steps_per_check: Lanczos steps the batched solver queues between
            reads of the per-chain statuses; each chain's convergence test
            runs on the device and a converged chain's result is frozen
            until the read (1 = read after every step).
        grouped: sweep all length groups in lockstep with one launch per
            kernel per step for all chains whose current site has the same
            shape (default). Rightward sweeps align the groups at the left
            end, leftward sweeps at the right end, so the site shapes agree
            everywhere except within log2(chi) sites of the far end of the
            shorter chains; a step takes one launch set per distinct shape.
        concurrent: with ``grouped=False``, sweep the length groups
            concurrently (one host thread and stream each) instead of one
            after another. Off by default: that form has ended in a host
            fault inside the launch path under a kernel tracer, a cause not
            yet found.
    """

    def __init__(
        self,
        mpos: Sequence,
        bond_dim: Optional[int] = None,
        mps: Optional[Sequence] = None,
        seeds: Optional[Sequence[int]] = None,
        exact_solver_dim: int = 200,
        *,
        chi: Optional[int] = None,
        max_basis: int = 64,
        keep: int = 20,
        max_matvecs: int = 5000,
        steps_per_check: int = 8,
        grouped: bool = True,
        concurrent: bool = False,
        device: Optional[torch.device] = None,
    ):
        if chi is not None and bond_dim is not None and chi != bond_dim:
            raise ValueError("chi and bond_dim disagree")
        bond = chi if chi is not None else bond_dim
        if bond is None:
            raise TypeError("BatchedFiniteDMRG needs a bond dimension (bond_dim or chi)")
        if len(mpos) == 0:
            raise ValueError("no chains")
        self.device = device if device is not None else _lib.require_gpu()
        _lib.load_library()
        self._mpos = [as_mpo(m) for m in mpos]
        d = self._mpos[0].phys_dim
        if any(m.phys_dim != d for m in self._mpos):
            raise ValueError("all chains need the same physical dimension")
        self._bond_dim = int(bond)
        self._phys_dim = d
        self._exact_solver_dim = exact_solver_dim
        self._max_basis = max_basis
        self._keep = keep
        self._max_matvecs = max_matvecs
        self._steps_per_check = steps_per_check
        self._grouped = grouped
        self._concurrent = concurrent
        self._ws = _lib.Workspace(self.device)
        self._lock = threading.Lock()
        self.stats = {"matvecs": 0, "lockstep_matvecs": 0, "restarts": 0, "dense_sites": 0, "krylov_sites": 0}
        nchain = len(self._mpos)
        if mps is not None and len(mps) != nchain:
            raise ValueError("need one initial state per chain")
        if seeds is None:
            seeds = list(range(nchain))
        by_len: Dict[int, List[int]] = {}
        for i, m in enumerate(self._mpos):
            by_len.setdefault(m.n_sites, []).append(i)
        self._groups = []
        for n in sorted(by_len):
            idx = by_len[n]
            if mps is None:
                stacks = _random_stacks(n, self._bond_dim, d, [int(seeds[i]) for i in idx], self)
            else:
                starts = []
                for i in idx:
                    st = mps[i] if isinstance(mps[i], MatrixProductState) else MatrixProductState(mps[i], device=self.device)
                    if st.n_sites != n:
                        raise ValueError(f"chain {i}: MPO and MPS lengths differ")
                    starts.append(st.tensors)
                shapes = {tuple(tuple(t.shape) for t in s) for s in starts}
                if len(shapes) != 1:
                    raise ValueError(f"chains of length {n} have different bond dimensions")
                stacks = [torch.stack([s[k] for s in starts]).contiguous() for k in range(n)]
            self._groups.append(_Group(self, idx, [self._mpos[i] for i in idx], stacks))
        # the stream of grouped sweeps (default priority)
        self._gstream = torch.cuda.Stream(device=self.device)
        torch.cuda.synchronize(self.device)
        self._energies: List[List[float]] = [[] for _ in range(nchain)]
        self.sweep_times: List[float] = []
        # one worker thread per length group, kept for the life of this
        # object (created at the first concurrent sweep)
        self._pool: Optional[ThreadPoolExecutor] = None
        self._finalizer = None

    def close(self) -> None:
        """Stop the worker threads of concurrent sweeps."""
        if self._finalizer is not None:
            self._finalizer()
            self._finalizer = None
        self._pool = None

    def _workers(self) -> ThreadPoolExecutor:
        if self._pool is None:
            self._pool = ThreadPoolExecutor(max_workers=len(self._groups), thread_name_prefix="tnx-group")
            self._finalizer = weakref.finalize(self, self._pool.shutdown, wait=True)
        return self._pool

    def _stream(self) -> int:
        return _lib.stream_ptr(self.device)

    @property
    def n_chains(self) -> int:
        return len(self._mpos)

    @property
    def bond_dim(self) -> int:
        return self._bond_dim

    @property
    def groups(self) -> Dict[int, List[int]]:
        """Chain indices per chain length."""
        return {g.n: list(g.chains) for g in self._groups}

    def sweep(self, direction: Direction = Direction.RIGHTWARD, tol: float = 1e-8) -> np.ndarray:
        """One directional sweep of every chain; returns each chain's
        energy at the last site of the sweep. By default all length groups
        step in lockstep, one launch per kernel per step for every stack of
        equal site shapes; with ``grouped=False`` the groups sweep one after
        another or, with ``concurrent=True``, concurrently on one host
        thread and stream each (the library calls release the GIL)."""
        if self._grouped and len(self._groups) > 1:
            return self._sweep_grouped(direction, tol)
        out = np.empty(self.n_chains)
        if len(self._groups) == 1 or not self._concurrent:
            for g in self._groups:
                out[g.chains] = g.sweep(direction, tol)
            return out
        ex = self._workers()
        futs = [(g, ex.submit(g.sweep, direction, tol)) for g in self._groups]
        for g, f in futs:
            out[g.chains] = f.result()
        return out

    def _sweep_grouped(self, direction: Direction, tol: float) -> np.ndarray:
        """One directional sweep of all groups in lockstep: at step t each
        group with t < n - 1 is at site t (rightward) or n - 1 - t
        (leftward); the groups whose site and neighbour shapes agree form one
        stack, and each stack takes one launch per kernel for its solve,
        psi += 1e-5 H psi, gauge move and environment update. Per chain the
        arithmetic is that of the per-group sweep."""
        gs_all = self._groups
        ms = self._gstream
        right = direction == Direction.RIGHTWARD
        for g in gs_all:
            if not right:
                with torch.cuda.stream(g.stream):
                    g.ensure_left(g.n - 1)
            ms.wait_stream(g.stream)
        out = np.empty(self.n_chains)
        sp = ms.cuda_stream
        with torch.cuda.stream(ms):
            for t in range(max(g.n for g in gs_all) - 1):
                stacks: Dict[tuple, list] = {}
                for g in gs_all:
                    if t >= g.n - 1:
                        continue
                    site = t if right else g.n - 1 - t
                    nb = g.A[site + 1] if right else g.A[site - 1]
                    key = (g.dims(site), tuple(nb.shape[1:]))
                    stacks.setdefault(key, []).append((g, site))
                for (dims, _), members in stacks.items():
                    self._grouped_step(members, dims, right, tol, sp, out)
            ms.synchronize()
        return out

    def _grouped_step(self, members, dims, right: bool, tol: float, sp: int, out: np.ndarray) -> None:
        ws = self._ws
        B = sum(g.B for g, _ in members)
        L = _stack([g.left[s] for g, s in members])
        R = _stack([g.right[s] for g, s in members])
        W = _stack([g.W[s] for g, s in members])
        A0 = _stack([g.A[s] for g, s in members])
        chains = [c for g, _ in members for c in g.chains]
        sites = sorted({s for _, s in members})
        e, vec, hvec = _solve(self, ws, sp, L, W, R, A0, dims, B, chains, sites, tol)
        _lib.check(_lib.lib().tnx_axpy(vec.numel(), 1e-5, _lib.ptr(hvec, "hvec"), _lib.ptr(vec, "vec"), sp),
                   "tnx_axpy")
        A = vec.reshape(A0.shape)
        if right:
            nb = _gauge_right(ws, sp, A, _stack([g.A[s + 1] for g, s in members]), B)
            env = _env_left(ws, sp, L, A, W, dims, B)
        else:
            nb = _gauge_left(ws, sp, A, _stack([g.A[s - 1] for g, s in members]), B)
            env = _env_right(ws, sp, R, A, W, dims, B)
        off = 0
        for g, s in members:
            sl = slice(off, off + g.B)
            off += g.B
            g.A[s] = A[sl]
            if right:
                g.A[s + 1] = nb[sl]
                g.left[s + 1] = env[sl]
                last = s == g.n - 2
            else:
                g.A[s - 1] = nb[sl]
                g.right[s - 1] = env[sl]
                last = s == 1
            if last:
                out[g.chains] = e[sl]

    def run(self, tol: float = 1e-8, max_sweep: int = 100) -> List[List[float]]:
        """Sweep all chains (alternating directions) until every chain's
        energy changes by less than ``tol`` between sweeps or ``max_sweep``
        is reached (the stopping rule of ``tnpy/finite_dmrg.py:173-212`` per
        chain). Returns the per-sweep energies of every chain."""
        done = np.zeros(self.n_chains, dtype=bool)
        prev = np.full(self.n_chains, np.nan)
        for n_sweep, direction in zip(range(1, max_sweep + 1), cycle([Direction.RIGHTWARD, Direction.LEFTWARD])):
            t0 = time.perf_counter()
            e = self.sweep(direction, tol)
            torch.cuda.synchronize(self.device)
            self.sweep_times.append(time.perf_counter() - t0)
            for i in range(self.n_chains):
                if not done[i]:
                    self._energies[i].append(float(e[i]))
            grad = np.abs(e - prev)
            done |= grad < tol
            prev = e
            logger.info(
                f"Batched sweep {n_sweep}/{max_sweep}: {int(done.sum())}/{self.n_chains} chains converged, "
                f"{self.sweep_times[-1]:.3f} s"
            )
            if done.all():
                break
        return self.energies

    update = run

    @property
    def energies(self) -> List[List[float]]:
        return [list(e) for e in self._energies]

    def mps(self, chain: int) -> MatrixProductState:
        """The current state of one chain as a ``MatrixProductState``."""
        for g in self._groups:
            if chain in g.chains:
                k = g.chains.index(chain)
                return MatrixProductState([a[k].clone() for a in g.A], device=self.device)
        raise IndexError(chain)


def _random_stacks(n: int, chi: int, d: int, seeds: Sequence[int], owner: "BatchedFiniteDMRG"):
    """Stacked random initial states of equal-length chains: chain b's
    entries are those of ``MatrixProductState.random(n, chi, d, seeds[b])``
    (N(0,1) from ``numpy.random.default_rng(seed)``, site by site in C
    order), then one batched LQ sweep from the right end with the same
    per-site rescaling and final normalisation."""
    dims = [1] + _bond_dims(n, chi, d) + [1]
    shapes = [(dims[i], d, dims[i + 1]) for i in range(n)]
    per = [[] for _ in range(n)]
    for sd in seeds:
        rng = np.random.default_rng(sd)
        for k, sh in enumerate(shapes):
            per[k].append(rng.standard_normal(sh))
    A = [torch.as_tensor(np.stack(p), device=owner.device).contiguous() for p in per]
    B = len(seeds)
    lib = _lib.lib()
    for site in range(n - 1, 0, -1):
        _, l, p, r = A[site].shape
        prv = A[site - 1]
        out = torch.empty_like(prv)
        wp, wb = owner._ws.get(lib.tnx_gauge_batched_workspace_bytes(l, p, r, B))
        _lib.check(
            lib.tnx_gauge_left_batched(
                _lib.ptr(A[site], "site"), _lib.ptr(prv, "prev"), _lib.ptr(out, "prev_out"),
                l, p, r, prv.shape[1] * prv.shape[2], B, wp, wb, owner._stream(),
            ),
            "tnx_gauge_left_batched",
        )
        A[site - 1] = (out / torch.linalg.vector_norm(out, dim=(1, 2, 3), keepdim=True)).contiguous()
    return A


def random_coupling_chains(n_chains: int, lengths: Sequence[int], seed: int):
    """Chain lengths drawn uniformly from ``lengths`` and per-chain coupling
    seeds, all from ``numpy.random.default_rng(seed)``; returns
    (lengths, coupling seeds) for ``RandomCouplingHeisenberg``."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(lengths), size=n_chains)
    seeds = rng.integers(0, 2**31 - 1, size=n_chains)
    return [int(n) for n in ns], [int(s) for s in seeds]


__all__ = ["BatchedFiniteDMRG", "random_coupling_chains"]
