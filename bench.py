"""Benchmark: one-site finite DMRG sweeps at chi=1024 on one MI355X.

Workload (the chi=1024 configuration of the headline metric): XXZ chain,
n=200, delta=1.0 (tnpy's sign convention: the ferromagnetic Heisenberg
chain), MPO bond D=5, d=2, one-site DMRG at chi=1024, eigen tolerance 1e-8.
A "step" is one directional sweep (n-1 site updates: Krylov eigensolve,
perturbation, QR gauge move, environment update), as
``tnpy/finite_dmrg.py:143-171``.

The whole run keeps to a wall-clock budget (``--budget`` seconds after the
library is loaded), so the number of timed sweeps can be below ``--steps``;
``steps`` in the output line is the number actually timed in each
precision. The order:

1. Kernel rates at chi (HIP events on the library's stream): the H_eff
   matvec and its stage-1 GEMM in FP64 and the mixed arithmetic (the
   ``roofline`` object), and the QR gauge move.
2. Warm state, untimed: one-site DMRG at ``--warm-chi`` from a seeded
   random MPS until the energy changes by less than tol (a bond-dimension
   ramp), then the bonds are widened to chi (zero padding, Householder LQ
   sweep: every chi=1024 tensor is a full orthonormal basis) and ``--warmup``
   FP64 sweeps run at chi. The reference's own sweep clock is the mean of
   the fastest sweeps of a run (``tnpy/finite_dmrg.py:252-256``): the
   converged end of a run, which this reproduces.
3. Timed: the same number of sweeps in FP64 and in the mixed precision, each
   from the same snapshot of the warm state (``value`` = mixed sweeps/s).
4. Krylov sample: the convergence phase of a run (the first sweep from a
   random state), where the eigensolver dominates: ``--krylov-sites``
   updates at the middle of the chain from a seeded random chi=1024 MPS, in
   FP64 and in the mixed precision from the same state. Its first site is
   the identical local eigenproblem in both, so its |dE| measures the mixed
   solver's accuracy at chi=1024.
5. CPU baseline: the numpy FP64 restatement of the same site update (test
   oracle, scipy/OpenBLAS on the host cores) on a bounded sample, scaled to
   a sweep with the GPU's matvec count per site.
6. Many-chain leg (config of 256 random-coupling chains at chi=64), within
   what is left of the budget.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "DMRG sweeps/s & matvec TFLOP/s @χ=1024: mixed vs FP64, |ΔE|<1e-10"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2516.6 / 3.0}
HBM_GBS = 8000.0
_T0 = time.perf_counter()


def log(msg: str) -> None:
    print(f"[bench {time.perf_counter() - _T0:7.1f}s] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--budget", type=float, default=300.0,
                   help="wall-clock seconds for the whole run after the library is loaded")
    p.add_argument("--n", type=int, default=200)
    p.add_argument("--chi", type=int, default=1024)
    p.add_argument("--delta", type=float, default=1.0)
    p.add_argument("--tol", type=float, default=1e-8)
    p.add_argument("--seed", type=int, default=2024)
    p.add_argument("--mixed", default="fp32", choices=["fp32", "bf16x3"])
    p.add_argument("--max-basis", type=int, default=64)
    p.add_argument("--keep", type=int, default=20)
    p.add_argument("--warm-chi", type=int, default=16,
                   help="bond dimension of the ramp that prepares the warm state")
    p.add_argument("--warm-sweeps", type=int, default=30, help="sweep limit of the ramp")
    p.add_argument("--krylov-sites", type=int, default=6,
                   help="site updates of the Krylov sample (0 skips it)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--cpu-reps", type=int, default=2)
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--batch-chains", type=int, default=256)
    p.add_argument("--batch-chi", type=int, default=64)
    p.add_argument("--batch-lengths", default="32,48,64,96")
    p.add_argument("--batch-seed", type=int, default=7)
    p.add_argument("--batch-sweeps", type=int, default=2)
    p.add_argument("--only-batched", action="store_true")
    p.add_argument("--batched-mode", choices=["grouped", "serial", "concurrent"], default="grouped",
                   help="grouped: all length groups in lockstep, one launch per kernel per step for "
                        "each stack of equal site shapes; serial: groups one after another; "
                        "concurrent: groups on one host thread and stream each")
    p.add_argument("--no-batched-compare", action="store_true",
                   help="skip timing the per-group (serial) form beside the grouped one")
    p.add_argument("--parity", action="store_true",
                   help="also run the converged mixed-vs-FP64 comparison from a random state at chi "
                        "(n=--parity-n, --parity-sweeps sweeps each)")
    p.add_argument("--parity-n", type=int, default=24)
    p.add_argument("--parity-sweeps", type=int, default=6)
    p.add_argument("--two-site", type=int, default=0, metavar="K",
                   help="only time K two-site bond updates at the middle of a random-field Heisenberg "
                        "chain (n=--two-site-n, h=0.5, chi=--chi) in FP64 and in the mixed precision")
    p.add_argument("--two-site-n", type=int, default=128)
    p.add_argument("--two-site-sample", type=int, default=2, metavar="K",
                   help="two-site updates per precision in the default run's config-5 sample (0 skips it)")
    p.add_argument("--tiles", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tnpy_amd",
                                                  "gemm_tiles_mi355x.txt"),
                   help="GEMM tile table to load (if it exists) so tile choices repeat across runs")
    p.add_argument("--save-tiles", default=None, help="write the tile table after the run")
    p.add_argument("--kernels-only", action="store_true",
                   help="only the HIP-event kernel rates at chi, plus a plain 256 MiB device copy "
                        "(the reference kernel for scaling HBM counters)")
    p.add_argument("--truncation", type=int, default=0, metavar="R",
                   help="only the two-site truncation: R timed SVDs of a (2 chi) x (2 chi) block "
                        "with a known spectrum cut to chi")
    return p.parse_args()


class Budget:
    """Wall-clock budget shared by the legs of one run."""

    def __init__(self, seconds: float):
        self.end = time.perf_counter() + seconds

    def left(self) -> float:
        return self.end - time.perf_counter()


def copy_reference(mib=256, reps=10):
    """A plain device-to-device copy of known size (torch's copy kernel):
    reads and writes mib MiB per launch."""
    import torch

    x = torch.ones(mib * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    ms = event_time(lambda: y.copy_(x), reps)
    return {"bytes_read": x.numel() * 8, "bytes_written": y.numel() * 8, "ms": ms,
            "GBps": 2 * x.numel() * 8 / ms / 1e6}


def event_time(fn, reps):
    """Average ms per call of fn over reps calls, HIP events on torch's
    current stream (the stream every libtnx call here is issued on)."""
    import torch

    fn()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def kernel_rates(chi, D, d, precisions):
    """Matvec, dominant-GEMM and gauge-move rates at (chi, D, d)."""
    import torch

    from tnpy_amd import _lib

    L = _lib.lib()
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    Lt = torch.randn(chi, D, chi, dtype=torch.float64, device=dev, generator=g)
    Rt = torch.randn(chi, D, chi, dtype=torch.float64, device=dev, generator=g)
    W = torch.randn(D, D, d, d, dtype=torch.float64, device=dev, generator=g)
    x = torch.randn(chi, d, chi, dtype=torch.float64, device=dev, generator=g)
    y = torch.empty_like(x)
    ws = torch.empty(L.tnx_heff1_workspace_bytes(chi, D, D, d, chi), dtype=torch.uint8, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    F1 = 4 * D * d * chi**3 + 2 * D * D * d * d * chi * chi
    A = torch.randn(chi * d, chi, dtype=torch.float64, device=dev, generator=g)
    B = torch.randn(chi, D * chi, dtype=torch.float64, device=dev, generator=g)
    Cm = torch.empty(chi * d, D * chi, dtype=torch.float64, device=dev)
    FG = 2 * (chi * d) * (D * chi) * chi
    out = {}
    for p in precisions:
        pi = _lib.PRECISIONS[p]
        tm = event_time(
            lambda: L.tnx_heff1_matvec(Lt.data_ptr(), W.data_ptr(), Rt.data_ptr(), x.data_ptr(), y.data_ptr(),
                                       chi, D, D, d, chi, pi, ws.data_ptr(), ws.numel(), s), 10)
        tg = event_time(
            lambda: L.tnx_gemm_f64(chi * d, D * chi, chi, 1, 1.0, A.data_ptr(), chi, 0, 1, B.data_ptr(), D * chi,
                                   0, 1, 0.0, Cm.data_ptr(), D * chi, 0, pi, s), 20)
        out[p] = {
            "matvec_ms": tm,
            "matvec_tflops": F1 / tm / 1e9,
            "gemm_ms": tg,
            "gemm_tflops": FG / tg / 1e9,
            "gemm_flops_per_launch": FG,
        }
        if p != "fp64":
            # the form the low-precision Lanczos cycles run: FP32 basis vector
            # and FP32 environment copy staged directly (FP64 output)
            A32, B32 = A.float(), B.float()
            tg32 = event_time(
                lambda: L.tnx_gemm(chi * d, D * chi, chi, 1, 1.0, A32.data_ptr(), 1, chi, 0, 1, B32.data_ptr(), 1,
                                   D * chi, 0, 1, 0.0, Cm.data_ptr(), 0, D * chi, 0, pi, s), 20)
            out[p]["gemm_f32_operands_ms"] = tg32
            out[p]["gemm_f32_operands_tflops"] = FG / tg32 / 1e9
    # gauge move (QR of the chi*d x chi site matrix, R times the next site)
    site = torch.randn(chi, d, chi, dtype=torch.float64, device=dev, generator=g)
    work = site.clone()
    nxt = torch.randn(chi, d, chi, dtype=torch.float64, device=dev, generator=g)
    nout = torch.empty_like(nxt)
    gws = torch.empty(L.tnx_gauge_workspace_bytes(chi, d, chi), dtype=torch.uint8, device=dev)

    def gauge():
        work.copy_(site)
        L.tnx_gauge_right(work.data_ptr(), nxt.data_ptr(), nout.data_ptr(), chi, d, chi, d * chi,
                          gws.data_ptr(), gws.numel(), s)

    tq = event_time(gauge, 5)
    m, n = chi * d, chi
    fq = 2 * (2 * m * n * n - 2 * n**3 / 3) + 2 * n * n * (d * chi)  # QR + thin Q + R*next
    out["gauge"] = {"ms": tq, "tflops": fq / tq / 1e9, "flops": fq}
    return out


def measured_traffic(precision, M, N, K):
    """HBM bytes per launch of the stage-1 GEMM from the committed counter
    summary (rocprofv3 FETCH_SIZE / WRITE_SIZE runs, scaled by a plain
    copy's reading), or None when no summary matches this shape."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "gemm_stage1_chi1024_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except OSError:
        return None
    for t in data.get("entries", [data]):
        if (t.get("precision"), t.get("M"), t.get("N"), t.get("K")) == (precision, M, N, K):
            return t["bytes_per_launch"]
    return None


def make_dmrg(args, mpo, state, prec, dev, chi):
    from tnpy_amd import FiniteDMRG

    return FiniteDMRG(mpo, bond_dim=chi, mps=state, precision=prec, max_basis=args.max_basis,
                      keep=args.keep, compute_variance=False, device=dev)


def timed_sweep(d, direction, tol, label):
    """One sweep; returns (seconds, energy, matvecs, fp64 matvecs)."""
    import torch

    n = d.n_sites
    t0 = time.perf_counter()

    def progress(site, energy):
        done = site + 1 if direction.value == 1 else n - site
        if done % 50 == 0:
            log(f"  {label}: {done}/{n - 1} sites, E={energy:.12f}, {time.perf_counter() - t0:.1f}s")

    d.on_site = progress
    mv0, mv640 = d.stats["matvecs"], d.stats["matvecs_f64"]
    e = d.sweep(direction, tol=tol)
    torch.cuda.synchronize()
    d.on_site = None
    return time.perf_counter() - t0, e, d.stats["matvecs"] - mv0, d.stats["matvecs_f64"] - mv640


def warm_state(args, mpo, dev, budget):
    """Bond-dimension ramp: DMRG at --warm-chi from the seeded random MPS
    until converged, then the bonds widened to chi."""
    import torch

    from tnpy_amd import Direction, MatrixProductState

    n, chi = args.n, args.chi
    t = time.perf_counter()
    mps = MatrixProductState.random(n=n, bond_dim=args.warm_chi, phys_dim=2, seed=args.seed, device=dev)
    d = make_dmrg(args, mpo, mps, "fp64", dev, args.warm_chi)
    dirs = [Direction.RIGHTWARD, Direction.LEFTWARD]
    energies = []
    for k in range(args.warm_sweeps):
        e = d.sweep(dirs[k % 2], tol=args.tol)
        energies.append(e)
        if len(energies) >= 2 and abs(energies[-1] - energies[-2]) < args.tol and k % 2 == 1:
            break
        if budget.left() < 0.8 * budget.total:
            break
    torch.cuda.synchronize()
    ramp = {"chi": args.warm_chi, "sweeps": len(energies), "energy": energies[-1],
            "seconds": time.perf_counter() - t}
    log(f"ramp at chi={args.warm_chi}: {len(energies)} sweeps, E={energies[-1]:.14f}, {ramp['seconds']:.1f}s")
    # the ramp ends with a leftward sweep: centre at site 0, ready for a rightward sweep
    big = d.mps.grow(chi)
    del d
    torch.cuda.synchronize()
    log(f"state widened to bonds max {big.bond_dim}")
    return big, ramp


def krylov_sample(args, mpo, dev, budget):
    """The convergence phase: --krylov-sites rightward updates at the chain
    middle from a seeded random chi MPS (centre moved there by QR gauge
    moves, which leave the state unchanged), FP64 and the mixed precision
    from the same state."""
    import torch

    from tnpy_amd import Direction, MatrixProductState, _lib

    n, chi, K = args.n, args.chi, args.krylov_sites
    mps = MatrixProductState.random(n=n, bond_dim=chi, phys_dim=2, seed=args.seed + 1, device=dev, host_rng=False)
    mid = max(1, n // 2 - K // 2)
    ws = _lib.Workspace(dev)
    for site in range(mid):
        mps.split_tensor(site, Direction.RIGHTWARD, workspace=ws)
    snap = [t.clone() for t in mps.tensors]
    del mps, ws
    torch.cuda.synchronize()
    out = {}
    for prec in ("fp64", args.mixed):
        d = make_dmrg(args, mpo, MatrixProductState([t.clone() for t in snap], device=dev), prec, dev, chi)
        d.environment.ensure_left(mid)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        energies = []
        for site in range(mid, min(mid + K, n - 1)):
            energies.append(d._site_step(site, Direction.RIGHTWARD, args.tol))
            if budget.left() < 0:
                break
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[prec] = {"seconds": dt, "sites": len(energies), "s_per_site": dt / len(energies),
                     "matvecs": d.stats["matvecs"], "matvecs_f64": d.stats["matvecs_f64"], "energies": energies}
        log(f"krylov sample {prec}: {len(energies)} sites in {dt:.2f}s, matvecs {d.stats['matvecs']} "
            f"(f64 {d.stats['matvecs_f64']}), E={energies[-1]:.12f}")
        del d
        torch.cuda.empty_cache()
    e64, emx = out["fp64"]["energies"], out[args.mixed]["energies"]
    k = min(len(e64), len(emx))
    return {
        "workload": (f"XXZ n={n} delta={args.delta} chi={chi}: {k} one-site updates at sites {mid}..{mid + k - 1} "
                     f"of the first rightward sweep from a seeded random MPS, tol={args.tol}"),
        "fp64": {q: v for q, v in out["fp64"].items() if q != "energies"},
        "mixed": {q: v for q, v in out[args.mixed].items() if q != "energies"},
        "abs_dE_first_site": abs(e64[0] - emx[0]),
        "rel_dE_first_site": abs(e64[0] - emx[0]) / abs(e64[0]),
        "speedup_vs_fp64": (out["fp64"]["s_per_site"] / out[args.mixed]["s_per_site"]),
    }


def cpu_baseline(chi, D, d, n, mv_per_site, bond, reps):
    """Bounded CPU sample: the oracle's matvec, SVD gauge move and
    environment update at chi x chi, each timed ``reps`` times on the host
    cores, scaled to one sweep of the same chain."""
    from oracle import dmrg_ref as o

    rng = np.random.default_rng(0)
    Lt = rng.standard_normal((chi, D, chi))
    Rt = rng.standard_normal((chi, D, chi))
    W = o.xxz_bulk(1.0)
    x = rng.standard_normal((chi, d, chi))

    def best(fn):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t)
        return min(ts)

    t_mv = best(lambda: o.heff_matvec(Lt, W, Rt, x))
    arr = [x.copy(), rng.standard_normal((chi, d, chi))]
    t_svd = best(lambda: o.split_tensor_svd([a.copy() for a in arr], 0, o.RIGHTWARD))
    t_env = best(lambda: o.env_left(Lt, x, W))
    # scale per site by the bond dimensions of the chain (cost ~ chil*chir*(chil+chir))
    dims = [1] + bond + [1]
    ref = 2.0 * chi**3
    total = 0.0
    for i in range(n - 1):
        l, r = dims[i], dims[i + 1]
        f = l * r * (l + r) / ref
        total += f * ((mv_per_site + 1.0) * t_mv + t_env) + (l * d * r * r) / (d * chi**3) * t_svd
    return {"sweep_s": total, "t_matvec_s": t_mv, "t_svd_s": t_svd, "t_env_s": t_env}


def truncation_leg(chi, reps, dev):
    """The two-site truncation alone: a (2 chi) x (2 chi) block with singular
    values logspace(0, -12) between seeded random orthogonal factors, cut to
    chi by the library SVD, timed with HIP events on the library's stream;
    the kept values against the known spectrum."""
    import numpy as np
    import torch

    from tnpy_amd import _lib

    n, k = 2 * chi, chi
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    Ua, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    Vb, _ = torch.linalg.qr(torch.randn(n, n, dtype=torch.float64, device=dev, generator=g))
    s_true = np.logspace(0, -12, n)
    X = (Ua * torch.as_tensor(s_true, device=dev)) @ Vb.T
    ws = _lib.Workspace(dev)
    _lib.svd_trunc(X, k, ws)  # workspace and tile set-up outside the timed calls
    torch.cuda.synchronize()
    times, err = [], 0.0
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _, _, sv = _lib.svd_trunc(X, k, ws)
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
        err = max(err, float(np.abs(sv - s_true[:k]).max()))
    log(f"truncation {n}x{n} -> {k}: best {min(times):.1f} ms, mean {np.mean(times):.1f} ms, "
        f"max |s - s_true| {err:.2e}")
    return {"block": [n, n], "keep": k, "reps": reps, "best_ms": min(times), "mean_ms": float(np.mean(times)),
            "max_abs_sv_err": err}


def two_site_leg(args, dev):
    """Two-site DMRG sample: random-field Heisenberg chain (Delta=1, h_i ~
    U(-0.5, 0.5) from the seed), seeded random MPS at chi, orthogonality
    centre moved to the middle by QR gauge moves (the state is unchanged),
    then K rightward two-site updates (two-site Lanczos, truncation of the
    (chi*d) x (d*chi) block back to chi, environment update) timed in FP64
    and in the mixed precision from the same state."""
    import torch

    from tnpy_amd import Direction, FiniteDMRG, MatrixProductState, _lib
    from tnpy_amd.model import RandomHeisenberg

    n, chi, K = args.two_site_n, args.chi, args.two_site
    model = RandomHeisenberg(n=n, h=0.5, seed=args.seed)
    mps = MatrixProductState.random(n=n, bond_dim=chi, phys_dim=2, seed=args.seed, device=dev, host_rng=False)
    mid = max(0, n // 2 - K // 2)
    ws = _lib.Workspace(dev)
    for site in range(mid):
        mps.split_tensor(site, Direction.RIGHTWARD, workspace=ws)
    snap = [t.clone() for t in mps.tensors]
    del mps
    torch.cuda.synchronize()
    out = {}
    for prec in ("fp64", args.mixed):
        d = FiniteDMRG(model.mpo, bond_dim=chi, block_size=2, mps=MatrixProductState([t.clone() for t in snap], device=dev),
                       precision=prec, max_basis=args.max_basis, keep=args.keep, compute_variance=False, device=dev)
        d.environment.ensure_left(mid)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        energies = []
        for site in range(mid, min(mid + K, n - 1)):
            energies.append(d._two_site_step(site, Direction.RIGHTWARD, args.tol))
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[prec] = {"seconds": dt, "bonds": len(energies), "s_per_bond": dt / max(1, len(energies)),
                     "matvecs": d.stats["matvecs"], "matvecs_f64": d.stats["matvecs_f64"],
                     "energies": energies}
        log(f"two-site {prec}: {len(energies)} bonds in {dt:.2f}s ({dt / max(1, len(energies)):.3f} s/bond), "
            f"matvecs {d.stats['matvecs']} (f64 {d.stats['matvecs_f64']}), E={energies[-1]:.12f}")
        del d
        torch.cuda.empty_cache()
    # the truncation alone: the library SVD of a (chi*d) x (d*chi) block cut to chi
    theta = torch.randn(chi * 2, 2 * chi, dtype=torch.float64, device=dev)
    _lib.svd_trunc(theta, chi)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, _sv = _lib.svd_trunc(theta, chi)
    torch.cuda.synchronize()
    t_svd = time.perf_counter() - t0
    e64, emx = out["fp64"]["energies"], out[args.mixed]["energies"]
    return {
        "truncation_ms": 1e3 * t_svd,
        "workload": (f"random-field Heisenberg n={n} h=0.5 chi={chi}, {K} two-site updates at the chain middle "
                     f"(sites {mid}..{mid + K - 1}), tol={args.tol}"),
        "fp64": {k: v for k, v in out["fp64"].items() if k != "energies"},
        "mixed": {k: v for k, v in out[args.mixed].items() if k != "energies"},
        "precision": args.mixed,
        "abs_dE_first_bond": abs(e64[0] - emx[0]),
        "abs_dE_last_bond": abs(e64[-1] - emx[-1]),
        "speedup_vs_fp64": out["fp64"]["seconds"] / out[args.mixed]["seconds"],
    }


def batched_leg(args, dev, budget=None):
    """Many-chain workload: random-coupling Heisenberg chains (J_i ~ U(0,1)
    per bond from a per-chain seed), lengths drawn uniformly from
    --batch-lengths, chi=--batch-chi, FP64, all swept together by
    BatchedFiniteDMRG. The same number of sweeps is then run chain by chain
    with FiniteDMRG on one sample chain per length, from the same initial
    states: their energies must agree, and their times scaled by the number
    of chains of each length give the unbatched FP64 time."""
    import torch

    from tnpy_amd import Direction, FiniteDMRG
    from tnpy_amd.batched_dmrg import BatchedFiniteDMRG, random_coupling_chains
    from tnpy_amd.model import RandomCouplingHeisenberg

    lengths = [int(x) for x in args.batch_lengths.split(",")]
    ns, cseeds = random_coupling_chains(args.batch_chains, lengths, args.batch_seed)
    models = [RandomCouplingHeisenberg(n=n, seed=sd) for n, sd in zip(ns, cseeds)]
    chi = args.batch_chi
    t = time.perf_counter()
    seeds = [args.batch_seed * 100003 + i for i in range(len(ns))]
    mode = args.batched_mode

    def make(m):
        return BatchedFiniteDMRG([x.mpo for x in models], chi=chi, seeds=seeds, max_basis=args.max_basis,
                                 keep=args.keep, grouped=m == "grouped", concurrent=m == "concurrent",
                                 device=dev)

    bd = make(mode)
    torch.cuda.synchronize()
    log(f"batched: {len(ns)} chains {dict(sorted((n, ns.count(n)) for n in set(ns)))} chi={chi} "
        f"ready in {time.perf_counter() - t:.1f}s")
    sample = [bd.groups[n][0] for n in sorted(bd.groups)]
    starts = {i: bd.mps(i) for i in sample}
    directions = [Direction.RIGHTWARD, Direction.LEFTWARD]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eb = []
    sweeps = max(1, args.batch_sweeps)
    for k in range(sweeps):
        eb.append(bd.sweep(directions[k % 2], tol=args.tol))
        torch.cuda.synchronize()
        log(f"batched sweep {k + 1}: {time.perf_counter() - t0:.2f}s")
        if budget is not None and budget.left() < 3.0 * (time.perf_counter() - t0) / (k + 1):
            sweeps = k + 1
            break
    dt = time.perf_counter() - t0
    solves = bd.stats["dense_sites"] + bd.stats["krylov_sites"]
    compare = None
    if mode == "grouped" and not args.no_batched_compare and (budget is None or budget.left() > 2.5 * dt):
        # the same sweeps with the length groups one after another
        bs = make("serial")
        torch.cuda.synchronize()
        ts = time.perf_counter()
        es = [bs.sweep(directions[k % 2], tol=args.tol) for k in range(sweeps)]
        torch.cuda.synchronize()
        ds = time.perf_counter() - ts
        compare = {
            "mode": "serial",
            "seconds": ds,
            "ms_per_step": 1e3 * ds / sweeps,
            "solve_launch_sets": bs.stats["dense_sites"] + bs.stats["krylov_sites"],
            "lockstep_matvecs": bs.stats["lockstep_matvecs"],
            "grouped_speedup": ds / dt,
            "max_rel_dE_vs_grouped": float(np.max(np.abs(es[-1] - eb[-1]) / np.abs(eb[-1]))),
        }
        log(f"batched per-group form: {sweeps} sweep(s) in {ds:.2f}s ({compare['solve_launch_sets']} solve "
            f"launch sets against {solves} grouped); grouped {ds / dt:.2f}x faster; max rel |dE| "
            f"{compare['max_rel_dE_vs_grouped']:.2e}")
    single = {}
    dEmax = 0.0
    for i in sample:
        f = FiniteDMRG(models[i].mpo, bond_dim=chi, mps=starts[i], max_basis=args.max_basis, keep=args.keep,
                       compute_variance=False, device=dev)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        es = [f.sweep(directions[k % 2], tol=args.tol) for k in range(sweeps)]
        torch.cuda.synchronize()
        single[ns[i]] = time.perf_counter() - ts
        dEmax = max(dEmax, abs(es[-1] - eb[-1][i]) / abs(es[-1]))
    est = sum(single[n] for n in ns)
    log(f"batched: {sweeps} sweep(s) of {len(ns)} chains in {dt:.2f}s; unbatched estimate {est:.1f}s "
        f"(per-length single-chain sweep times {single}); max rel |dE| batched vs single {dEmax:.2e}")
    return {
        "workload": (f"{len(ns)} random-coupling Heisenberg chains, J_i~U(0,1), n in {lengths}, chi={chi}, "
                     f"one-site DMRG sweeps from seeded random states, FP64, tol={args.tol}"),
        "chains": len(ns),
        "lengths": {str(n): ns.count(n) for n in sorted(set(ns))},
        "sweeps": sweeps,
        "mode": mode,
        "solve_launch_sets": solves,
        "seconds": dt,
        "sweeps_per_s": sweeps / dt,
        "chain_sweeps_per_s": len(ns) * sweeps / dt,
        "ms_per_step": 1e3 * dt / sweeps,
        "unbatched_fp64_est_seconds": est,
        "speedup_vs_unbatched_fp64": est / dt,
        "sample_rel_dE_vs_unbatched": dEmax,
        "matvecs": bd.stats["matvecs"],
        "lockstep_matvecs": bd.stats["lockstep_matvecs"],
        "per_group_form": compare,
    }


def converged_parity(args, dev):
    """Energy agreement of the two arithmetics at convergence from a random
    state: an XXZ chain of --parity-n sites (n=24 puts its bulk sites at the
    full chi=1024) swept --parity-sweeps times, alternating direction, from
    one seeded random MPS in FP64 and in the mixed precision."""
    import torch

    from tnpy_amd import Direction, MatrixProductState
    from tnpy_amd.model import XXZ

    n, chi = args.parity_n, args.chi
    mpo = XXZ(n=n, delta=args.delta).mpo
    mps = MatrixProductState.random(n=n, bond_dim=chi, phys_dim=2, seed=args.seed + 1, device=dev, host_rng=False)
    snap = [t.clone() for t in mps.tensors]
    del mps
    out = {"workload": f"XXZ chain n={n} delta={args.delta} chi={chi}, {args.parity_sweeps} one-site sweeps "
                       f"(tol={args.tol}) from one seeded random MPS in each arithmetic"}
    dirs = [Direction.RIGHTWARD, Direction.LEFTWARD]
    for prec in ("fp64", args.mixed):
        d = make_dmrg(args, mpo, MatrixProductState([t.clone() for t in snap], device=dev), prec, dev, chi)
        t0 = time.perf_counter()
        es = [d.sweep(dirs[k % 2], tol=args.tol) for k in range(args.parity_sweeps)]
        torch.cuda.synchronize()
        out[prec] = {"energies": es, "seconds": time.perf_counter() - t0, "matvecs": d.stats["matvecs"]}
        log(f"parity {prec}: energies {', '.join(f'{e:.14f}' for e in es)}")
        del d
        torch.cuda.empty_cache()
    out["abs_dE"] = abs(out[args.mixed]["energies"][-1] - out["fp64"]["energies"][-1])
    if args.delta == 1.0:
        out["exact"] = -(n - 1) / 4.0
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the DMRG sweep here is "
            "single-GPU; there is no multi-GPU path)",
            file=sys.stderr,
        )
        sys.exit(2)
    import logging

    import torch

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from tnpy_amd import Direction, MatrixProductState, _lib
    from tnpy_amd.model import XXZ

    logging.getLogger("tnpy").setLevel(logging.WARNING)
    dev = torch.device("cuda", 0)
    _lib.load_library()
    torch.cuda.synchronize()
    log("library loaded")
    if args.tiles and os.path.exists(args.tiles):
        log(f"GEMM tile table {args.tiles}: {_lib.lib().tnx_tune_load(args.tiles.encode())} entries")
    if args.save_tiles:
        import atexit

        atexit.register(lambda: log(f"saved {_lib.lib().tnx_tune_save(args.save_tiles.encode())} tile entries "
                                    f"to {args.save_tiles}"))
    budget = Budget(args.budget)
    budget.total = args.budget
    if args.only_batched:
        print(json.dumps({"batched": batched_leg(args, dev)}), flush=True)
        return
    if args.two_site > 0:
        print(json.dumps({"two_site": two_site_leg(args, dev)}), flush=True)
        return
    if args.truncation > 0:
        print(json.dumps({"truncation": truncation_leg(args.chi, args.truncation, dev)}), flush=True)
        return
    if args.kernels_only:
        # an untimed pass first (see below); both passes are printed
        first = kernel_rates(args.chi, 5, 2, ["fp64", "fp32", "bf16x3"])
        rates = kernel_rates(args.chi, 5, 2, ["fp64", "fp32", "bf16x3"])
        print(json.dumps({"chi": args.chi, "rates": rates, "first_pass": first, "copy": copy_reference()}),
              flush=True)
        return

    n, chi = args.n, args.chi
    # 1. kernel rates at chi (the third arithmetic too, for its roofline)
    other = "bf16x3" if args.mixed == "fp32" else "fp32"
    # One untimed pass first: in a fresh process the first launches on newly
    # allocated operands run slower (measured 3-10 % on the FP32 and BF16x3
    # GEMMs, even after seconds of other GPU work); the second pass, on the
    # same blocks from the allocator's cache, is the state the kernels run
    # in during the sweeps.
    kernel_rates(chi, 5, 2, ["fp64", args.mixed, other])
    rates = kernel_rates(chi, 5, 2, ["fp64", args.mixed, other])
    for p in ("fp64", args.mixed):
        r = rates[p]
        log(f"{p}: matvec {r['matvec_ms']:.3f} ms ({r['matvec_tflops']:.1f} TF/s), "
            f"stage-1 GEMM {r['gemm_ms']:.3f} ms ({r['gemm_tflops']:.1f} TF/s)")
    log(f"gauge move (QR {2 * chi}x{chi} + R*next): {rates['gauge']['ms']:.3f} ms")

    # 2. warm state
    mpo = XXZ(n=n, delta=args.delta).mpo
    state, ramp = warm_state(args, mpo, dev, budget)
    d64 = make_dmrg(args, mpo, state, "fp64", dev, chi)
    directions = [Direction.RIGHTWARD, Direction.LEFTWARD]
    warm = []
    for w in range(args.warmup):
        dt, e, mv, _ = timed_sweep(d64, directions[w % 2], args.tol, f"warmup {w + 1}")
        warm.append({"seconds": dt, "energy": e, "matvecs": mv})
        log(f"warmup sweep {w + 1}: {dt:.2f}s E={e:.14f} matvecs={mv}")
        if budget.left() < 0.55 * budget.total:
            break
    snap = [t.clone() for t in d64.mps.tensors]
    start_dir = len(warm) % 2
    bond = d64.mps.bond_dims()
    del d64, state
    torch.cuda.empty_cache()

    # 3. timed sweeps: FP64 within its share of the budget, then the same
    # number in the mixed precision
    results = {}
    K = args.steps
    for prec in ("fp64", args.mixed):
        d = make_dmrg(args, mpo, MatrixProductState([t.clone() for t in snap], device=dev), prec, dev, chi)
        if directions[start_dir] == Direction.LEFTWARD:
            d.environment.ensure_left(n - 1)  # environment set-up stays outside the timed sweep
        torch.cuda.synchronize()
        share_end = time.perf_counter() + 0.25 * budget.total
        per = []
        energies, mvs, mv64s = [], 0, 0
        for k in range(K):
            dt, e, mv, mv64 = timed_sweep(d, directions[(start_dir + k) % 2], args.tol, f"{prec} sweep {k + 1}")
            per.append(dt)
            energies.append(e)
            mvs += mv
            mv64s += mv64
            if prec == "fp64" and time.perf_counter() + dt > share_end:
                K = k + 1
                break
        torch.cuda.synchronize()
        total = float(sum(per))
        results[prec] = {
            "seconds": total,
            "sweeps": len(per),
            "sweeps_per_s": len(per) / total,
            "best3_s": float(np.mean(sorted(per)[:3])),
            "energies": energies,
            "matvecs": mvs,
            "matvecs_f64": mv64s,
            "krylov_sites": d.stats["krylov_sites"],
        }
        log(f"{prec}: {len(per)} sweep(s) in {total:.2f}s ({len(per) / total:.3f} sweeps/s), "
            f"E={energies[-1]:.14f}, matvecs={mvs} (f64 {mv64s})")
        del d
        torch.cuda.empty_cache()
    K = results["fp64"]["sweeps"]
    dE = abs(results[args.mixed]["energies"][-1] - results["fp64"]["energies"][-1])
    exact = -(n - 1) / 4.0 if args.delta == 1.0 else None

    # 4. Krylov sample (convergence phase)
    krylov = None
    if args.krylov_sites > 0 and budget.left() > 0.2 * budget.total:
        krylov = krylov_sample(args, mpo, dev, budget)

    # 4b. two-site sample (random-field Heisenberg, chi, truncation by the library SVD)
    two_site = None
    if args.two_site_sample > 0 and budget.left() > 0.2 * budget.total:
        args.two_site = args.two_site_sample
        two_site = two_site_leg(args, dev)

    # 5. CPU baseline
    cpu = None
    if not args.no_cpu:
        mv_site = results["fp64"]["matvecs"] / max(1, results["fp64"]["krylov_sites"])
        cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
        log(f"CPU baseline sample (numpy FP64 restatement, {cores} threads)")
        c = cpu_baseline(chi, 5, 2, n, mv_site, bond, args.cpu_reps)
        cpu = {
            "value": 1.0 / c["sweep_s"],
            "unit": "sweeps/s",
            "cores": cores,
            "kind": "port",
            "sample": (
                f"numpy/OpenBLAS FP64 oracle at chi={chi}: best of {args.cpu_reps} of one H_eff matvec "
                f"({c['t_matvec_s']:.3f}s), one SVD gauge move ({c['t_svd_s']:.3f}s), one environment update "
                f"({c['t_env_s']:.3f}s); scaled to a sweep of the timed workload with {mv_site:.2f} matvecs/site "
                f"(the GPU FP64 count) and each site's bond dimensions"
            ),
        }
        log(f"CPU baseline: {cpu['value']:.3e} sweeps/s ({c['sweep_s']:.1f} s/sweep est.)")

    # 6. many-chain leg
    batched = None
    if not args.no_batched and budget.left() > 30.0:
        batched = batched_leg(args, dev, budget)
    parity = converged_parity(args, dev) if args.parity else None

    mixed = results[args.mixed]
    r = rates[args.mixed]
    line = {
        "metric": METRIC,
        "value": mixed["sweeps_per_s"],
        "unit": "sweeps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": len(warm),
        "ms_per_step": 1e3 * mixed["seconds"] / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.mixed} Krylov matvecs + f64 residual, f64 env/QR",
        "data": "synthetic (seeded random MPS, XXZ MPO)",
        "config": {
            "workload": (f"XXZ chain n={n} delta={args.delta} chi={chi}, one-site DMRG sweeps of a converged run "
                         f"(warm state: DMRG at chi={args.warm_chi}, bonds widened to {chi}, {len(warm)} sweeps "
                         f"at {chi}), tol={args.tol}"),
            "n": n, "chi": chi, "delta": args.delta, "D": 5, "d": 2, "tol": args.tol, "seed": args.seed,
            "max_basis": args.max_basis, "keep": args.keep,
        },
        "fp64": {
            "sweeps_per_s": results["fp64"]["sweeps_per_s"],
            "ms_per_step": 1e3 * results["fp64"]["seconds"] / K,
            "best3_ms": 1e3 * results["fp64"]["best3_s"],
            "energy": results["fp64"]["energies"][-1],
            "matvecs": results["fp64"]["matvecs"],
            "matvec_tflops": rates["fp64"]["matvec_tflops"],
        },
        "mixed": {
            "precision": args.mixed,
            "best3_ms": 1e3 * mixed["best3_s"],
            "energy": mixed["energies"][-1],
            "abs_dE_vs_fp64": dE,
            "matvecs": mixed["matvecs"],
            "matvecs_f64": mixed["matvecs_f64"],
            "matvec_tflops": r["matvec_tflops"],
            "speedup_vs_fp64": mixed["sweeps_per_s"] / results["fp64"]["sweeps_per_s"],
        },
        "exact_energy": exact,
        "fp64_abs_err_vs_exact": None if exact is None else abs(results["fp64"]["energies"][-1] - exact),
        "warm": {"ramp": ramp, "sweeps": warm},
        "krylov_sample": krylov,
        "two_site_sample": two_site,
        "roofline": {
            "bound": "mfma",
            "kernel": f"gemm_kernel ({args.mixed}) H_eff stage 1, M={2 * chi} N={5 * chi} K={chi}",
            "achieved": r["gemm_tflops"],
            "peak": PEAK_TFLOPS[args.mixed],
            "unit": "TFLOP/s",
            "frac": r["gemm_tflops"] / PEAK_TFLOPS[args.mixed],
            "traffic": measured_traffic(args.mixed, 2 * chi, 5 * chi, chi),
        },
        "roofline_fp64": {
            "bound": "mfma",
            "kernel": f"gemm_kernel (fp64) H_eff stage 1, M={2 * chi} N={5 * chi} K={chi}",
            "achieved": rates["fp64"]["gemm_tflops"],
            "peak": PEAK_TFLOPS["fp64"],
            "unit": "TFLOP/s",
            "frac": rates["fp64"]["gemm_tflops"] / PEAK_TFLOPS["fp64"],
            "traffic": measured_traffic("fp64", 2 * chi, 5 * chi, chi),
        },
        "roofline_f32_operands": None if "gemm_f32_operands_tflops" not in r else {
            "bound": "mfma",
            "kernel": (f"gemm_kernel ({args.mixed}, FP32 operands as in the low-precision Lanczos cycles) "
                       f"H_eff stage 1, M={2 * chi} N={5 * chi} K={chi}"),
            "achieved": r["gemm_f32_operands_tflops"],
            "peak": PEAK_TFLOPS[args.mixed],
            "unit": "TFLOP/s",
            "frac": r["gemm_f32_operands_tflops"] / PEAK_TFLOPS[args.mixed],
            "traffic": None,
        },
        f"roofline_{other}": {
            "bound": "mfma",
            "kernel": f"gemm_kernel ({other}) H_eff stage 1, M={2 * chi} N={5 * chi} K={chi}",
            "achieved": rates[other]["gemm_tflops"],
            "peak": PEAK_TFLOPS[other],
            "unit": "TFLOP/s",
            "frac": rates[other]["gemm_tflops"] / PEAK_TFLOPS[other],
            "traffic": None,
        },
        "gauge_move": {"ms": rates["gauge"]["ms"], "tflops": rates["gauge"]["tflops"],
                       "frac_fp64_peak": rates["gauge"]["tflops"] / PEAK_TFLOPS["fp64"]},
        "cpu_baseline": cpu,
        "batched": batched,
        "parity": parity,
        "bench_seconds": time.perf_counter() - _T0,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
