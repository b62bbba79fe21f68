"""Benchmark of the two-site effective-Hamiltonian path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--chi 1024] [--no-cpu] [--sweep-steps S]

Workload: the spin-1/2 XXZ chain (Jxy = Jz = 1, MPO bond D = 5, d = 2), N = 100 sites, bond
dimension chi = 1024. A random MPS is brought to mixed-canonical form around the bulk bond
(sites 49, 50) and the environments L[49], R[51] are built with the library's own
environment kernels, so the local problem is a real projection of the chain Hamiltonian.

A "step" is one application of H_eff to theta (a = b = chi): the mixed-precision form (float32
operator on the float32 matrix cores, float64 Krylov vectors in and out). The float64 form of
the same matvec is timed beside it. The JSON line also carries:
  * fp64: the float64 matvec,
  * local_dE: |E_mixed - E_fp64| of the local ground state solved by the mixed two-stage
    Lanczos and by float64 Lanczos (residual 1e-9),
  * steps: bulk DMRG local steps (eigensolve + Gram truncation + environment), per step and
    extrapolated to a 198-step sweep, for both forms,
  * roofline: the slowest kernel of the mixed matvec (events on the launch stream),
  * cpu_baseline: the numpy float64 oracle matvec on the host cores (bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DMRG sweep s + H_eff TFLOP/s at chi=1024, mixed vs FP64 HIP @ dE<=1e-9"
PEAK = {"f64": 78.6, "f32": 157.3}          # MI355X dense matrix-core TFLOP/s
HBM_PEAK = 8000.0                           # GB/s


def two_site_flops(a, b, d1, d2, wl, wr):
    """4 d^2 D chi^3 + 2 d^4 D^2 chi^2 at a = b = chi (tnkit/bench/profile/gpu_matvec_prof.py:92)."""
    return (2 * a * wl * a * d1 * d2 * b + 2 * (d1 * d2 * wr) * (wl * d1 * d2) * a * b
            + 2 * a * d1 * d2 * wr * b * b)


def stage_flops(a, b, d1, d2, wl, wr):
    f = {1: 2 * a * wl * a * d1 * d2 * b, 2: 2 * (d1 * d2 * wr) * (wl * d1 * d2) * a * b,
         4: 2 * a * d1 * d2 * wr * b * b}
    f[3] = f[1] + f[2]      # GEMM1 with the MPO mix in its epilogue
    return f


def stage_bytes(a, b, d1, d2, wl, wr, e_op, e_io):
    """Algorithmic HBM bytes per stage: operands read once, result written once."""
    x1 = d1 * d2 * b * a * wl * e_op
    x2 = wr * b * a * d1 * d2 * e_op
    return {1: a * d1 * d2 * b * e_io + a * a * wl * e_op + x1,
            2: x1 + x2,
            3: a * d1 * d2 * b * e_io + a * a * wl * e_op + x2,
            4: x2 + wr * b * b * e_op + a * d1 * d2 * b * e_io}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "matvec_traffic.json")


def measured_traffic(kernel_label, chi, form):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the counter reads for a copy
    kernel of known size). None when no summary covers this kernel at this chi."""
    try:
        with open(TRAFFIC_FILE) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None, None
    best = None
    for row in table.get("rows", []):
        if row["chi"] == chi and row["form"] == form and kernel_label.startswith(row["stage"]):
            if best is None or len(row["stage"]) > len(best["stage"]):
                best = row
    if best is None:
        return None, None
    return best["fetch_bytes"] + best["write_bytes"], os.path.relpath(TRAFFIC_FILE, os.getcwd())


def event_time(fn, reps, stream):
    """Average ms of `fn` over `reps` calls, with HIP events on `stream`."""
    import torch
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    e.synchronize()
    return s.elapsed_time(e) / reps


def cgs2_bench(n, nlive=41, reps=20):
    """One Lanczos orthogonalisation (`kernels.krylov_orth`: the three-term subtraction, alpha,
    two classical Gram-Schmidt passes and beta) on a float64 basis of `nlive` rows of length n,
    HIP events on the launch stream. Algorithmic bytes per call: the pre-pass reads w, V[j] and
    V[j-1] and writes w (4 n 8 B); each of the three passes reads the nlive rows and w and writes
    w ((nlive + 2) n 8 B)."""
    import torch
    from tnkit_amd import kernels
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    V = torch.randn((nlive, n), dtype=torch.float64, device=dev, generator=g)
    V, _ = torch.linalg.qr(V.T)
    V = V.T.contiguous()
    w0 = torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    w = w0.clone()
    kw = kernels.KrylovWorkspace(n, dev)
    j = nlive - 1

    def call():
        w.copy_(w0)
        kernels.krylov_orth(V, nlive, w, j, [0.25], j - 1, kw)
    call()
    resid = float((V @ w).abs().max() / w.norm())
    stream = torch.cuda.current_stream()
    ms = event_time(call, reps, stream)
    ms_copy = event_time(lambda: w.copy_(w0), reps, stream)
    byts = (4 + 3 * (nlive + 2)) * n * 8
    t = ms - ms_copy
    return {"n": n, "nlive": nlive, "ms": t, "bytes": byts, "TBps": byts / (t * 1e-3) / 1e12,
            "orthogonality_after": resid,
            "note": "per call, less the timed refresh copy of w; HBM peak 8 TB/s, a flat "
                    "float64 read measured 6.4 TB/s on this part"}


def build_local_problem(N, chi, seed=0):
    """Bulk two-site problem of the XXZ chain from a random mixed-canonical MPS."""
    import numpy as np
    import torch
    from tnkit_amd import kernels
    from tnkit_amd.models import xxz_mpo

    dev = torch.device("cuda", 0)
    mpo = [torch.as_tensor(W).to(dev) for W in xxz_mpo(N)]
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    dims = [1] + [min(chi, 2 ** min(i, N - i)) for i in range(1, N)] + [1]
    i0 = N // 2 - 1
    M = [torch.randn((dims[i], 2, dims[i + 1]), generator=g, dtype=torch.float64, device=dev)
         for i in range(N)]
    for i in range(i0):                          # left-canonical left of the centre
        a, d, b = M[i].shape
        Q, Rr = torch.linalg.qr(M[i].reshape(a * d, b))
        M[i] = Q.reshape(a, d, -1).contiguous()
        M[i + 1] = torch.tensordot(Rr, M[i + 1], dims=([1], [0])).contiguous()
        M[i + 1] /= torch.linalg.norm(M[i + 1])
    for i in range(N - 1, i0 + 1, -1):            # right-canonical right of the pair
        a, d, b = M[i].shape
        Q, Rr = torch.linalg.qr(M[i].reshape(a, d * b).t())
        M[i] = Q.t().reshape(-1, d, b).contiguous()
        M[i - 1] = torch.tensordot(M[i - 1], Rr.t(), dims=([2], [0])).contiguous()
        M[i - 1] /= torch.linalg.norm(M[i - 1])
    L = torch.ones((1, 1, 1), dtype=torch.float64, device=dev)
    for i in range(i0):
        L = kernels.update_left_env(L, M[i], mpo[i])
    R = torch.ones((1, 1, 1), dtype=torch.float64, device=dev)
    for i in range(N - 1, i0 + 1, -1):
        R = kernels.update_right_env(R, M[i], mpo[i])
    theta = torch.tensordot(M[i0], M[i0 + 1], dims=([2], [0])).contiguous()
    theta /= torch.linalg.norm(theta)
    return dict(mpo=mpo, M=M, L=L, R=R, theta=theta, i0=i0)


def cpu_baseline(prob, budget_s=20.0):
    """numpy float64 oracle matvec on the host cores: a bounded sample of whole matvecs."""
    import numpy as np
    from oracle import dense as D
    try:
        from threadpoolctl import threadpool_info
        blas = [f"{i.get('internal_api')} {i.get('version')}" for i in threadpool_info()
                if i.get("user_api") == "blas"]
    except Exception:  # pragma: no cover
        blas = []
    L = prob["L"].cpu().numpy()
    R = prob["R"].cpu().numpy()
    W = prob["mpo"][prob["i0"]].cpu().numpy()
    th = prob["theta"].cpu().numpy()
    op = D.fused_two_site(L, W, W, R)
    a, d1, d2, b = th.shape
    fl = two_site_flops(a, b, d1, d2, W.shape[0], W.shape[1])
    D.apply_two_site_fused(th, *op)              # warm-up (page-in, BLAS threads)
    n, t0 = 0, time.perf_counter()
    while True:
        D.apply_two_site_fused(th, *op)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 10:
            break
    sec = el / n
    cores = int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))
    out = {"value": fl / sec / 1e12, "unit": "TFLOP/s", "cores": cores, "kind": "port",
           "sample": f"{n} float64 matvecs of the chi={a} XXZ bulk problem with the numpy oracle "
                     f"({sec * 1e3:.0f} ms each; BLAS: {', '.join(blas) or 'numpy default'})",
           "cpu_model": _cpu_model(), "blas": blas}
    try:                                          # the same matvec on one core
        from threadpoolctl import threadpool_limits
        with threadpool_limits(1):
            t1 = time.perf_counter()
            D.apply_two_site_fused(th, *op)
            s1 = time.perf_counter() - t1
        out["one_thread"] = {"value": fl / s1 / 1e12, "unit": "TFLOP/s", "cores": 1,
                             "sample": f"1 float64 matvec ({s1:.2f} s)"}
    except Exception as exc:  # pragma: no cover - threadpoolctl is in the image
        out["one_thread"] = {"error": str(exc)}
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def config1_cpu_vs_gpu():
    """Config 1 (TFIM N=32, g=1, chi=32, 10 sweeps, tol=rtol=0, seed 0, float64) in full on the
    host cores with the numpy oracle and on the device, each against the free-fermion energy."""
    import torch
    from oracle import dense as D
    from oracle import exact as X
    from tnkit_amd import _dmrg as H
    from tnkit_amd.models import tfim_mpo
    mpo = tfim_mpo(32, 1.0)
    kw = dict(n_sweeps=10, tol=0.0, rtol=0.0, seed=0)
    e_exact = X.tfim_free_fermion_energy(32, 1.0)
    t0 = time.perf_counter()
    e_cpu, _, _ = D.dmrg(mpo, 32, **kw)
    t_cpu = time.perf_counter() - t0
    H.dmrg(mpo, 32, n_sweeps=1, tol=0.0, rtol=0.0, seed=0, backend="hip-real")   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e_gpu, _, info = H.dmrg(mpo, 32, backend="hip-real", **kw)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    return {"workload": "TFIM N=32 g=1, two-site DMRG chi=32, 10 sweeps, tol=rtol=0, seed 0",
            "cpu_s": t_cpu, "gpu_s": t_gpu, "E_cpu": e_cpu, "E_gpu": e_gpu, "E_exact": e_exact,
            "dE_cpu": abs(e_cpu - e_exact), "dE_gpu": abs(e_gpu - e_exact),
            "gpu_matvecs": info["matvec_calls"],
            "note": "launch-bound at chi=32: the CPU is the faster device here"}


def local_solves(th, op64, op32, ws, chi, op16=None):
    """The local ground state by float64 thick-restart Lanczos and by the mixed two-stage form
    (float32 operator, then float64); with `op16`, also the two-stage form on the BF16 operator,
    whose stage 1 stops at a 3e-2 residual (the bf16 operator's own error is a few 1e-3)."""
    import torch
    from tnkit_amd import kernels
    from tnkit_amd.eigensolver import lanczos_ground_mixed, lanczos_ground_restart
    shape = tuple(th.shape)
    cnt = {"lo": 0, "hi": 0}
    mvh = lambda x: kernels.apply_two_site_fused(x.reshape(shape), *op64, ws=ws).reshape(-1)  # noqa
    mvl = lambda x: kernels.apply_two_site_fused(x.reshape(shape), *op32, ws=ws).reshape(-1)  # noqa
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    e64, _, i64 = lanczos_ground_restart(mvh, th.reshape(-1).clone(), tol=1e-9, max_iter=4 * chi)
    torch.cuda.synchronize()
    t_e64 = time.perf_counter() - t1
    t1 = time.perf_counter()
    emx, _, imx = lanczos_ground_mixed(mvl, mvh, th.reshape(-1).clone(), tol=1e-9,
                                       max_iter=4 * chi, counts=cnt)
    torch.cuda.synchronize()
    t_emx = time.perf_counter() - t1
    res = {"fp64_E": e64, "mixed_E": emx, "abs": abs(emx - e64), "fp64_matvecs": i64["iters"],
           "mixed_lo": cnt["lo"], "mixed_hi": cnt["hi"], "fp64_s": t_e64, "mixed_s": t_emx}
    if op16 is not None:
        c16 = {"lo": 0, "hi": 0}
        mv16 = lambda x: kernels.apply_two_site_fused(x.reshape(shape), *op16, ws=ws).reshape(-1)  # noqa
        t1 = time.perf_counter()
        e16, _, i16 = lanczos_ground_mixed(mv16, mvh, th.reshape(-1).clone(), tol=1e-9,
                                           max_iter=4 * chi, tol_lo=3e-2, counts=c16)
        torch.cuda.synchronize()
        res.update({"bf16_E": e16, "bf16_abs": abs(e16 - e64), "bf16_lo": c16["lo"],
                    "bf16_hi": c16["hi"], "bf16_stage1_E": i16["e_lo"],
                    "bf16_s": time.perf_counter() - t1})
    return res


def gaussian_sectors(chi, sigma, half_width):
    """Sector dimensions of one even-charge bond: charges 2k, |k| <= half_width, dims proportional
    to exp(-k^2 / (2 sigma^2)), at least 1, summing to chi (the largest ~ chi / (sigma sqrt(2 pi)))."""
    import numpy as np
    k = np.arange(-half_width, half_width + 1)
    w = np.exp(-k ** 2 / (2.0 * sigma ** 2))
    dims = np.maximum(1, np.floor(chi * w / w.sum())).astype(np.int64)
    while dims.sum() < chi:
        dims[np.argmax(chi * w / w.sum() - dims)] += 1
    while dims.sum() > chi:
        dims[np.argmax(dims)] -= 1
    return {int(2 * kk): int(d) for kk, d in zip(k, dims)}


def build_block_problem(chi=2048, sigma=2.7, half_width=15, seed=0):
    """The bulk two-site problem of the U(1) XXZ chain (2 S^z_total = 0) at bond chi split over
    2 half_width + 1 sectors of Gaussian sizes. The environments are those of a random block MPS
    whose bonds grow from each edge (bond n: the charges n sites can reach, dims the Gaussian
    profile capped by the number of basis states) over 2 half_width + 10 sites, built with the
    block environment kernels and rescaled by their largest entry after each site; so H_eff is
    symmetric, as a chain projection is."""
    import math
    import numpy as np
    import torch
    from tnkit_amd import blocks as B
    from tnkit_amd import block_ops as BO
    from tnkit_amd import block_mps as BM
    from tnkit_amd.models import xxz_mpo

    dev = torch.device("cuda", 0)
    target = gaussian_sectors(chi, sigma, half_width)
    peak = max(target.values())
    qmax = 2 * half_width
    nh = qmax + 10

    def leg(n):
        if n == nh:
            cs = sorted(target)
            return B.Leg(tuple(cs), tuple(target[c] for c in cs), 1)
        cs = [q for q in range(-min(n, qmax + 1), min(n, qmax + 1) + 1) if (q - n) % 2 == 0]
        dims = [min(max(1, int(round(peak * math.exp(-(q / 2.0) ** 2 / (2 * sigma ** 2))))),
                    math.comb(n, (n + q) // 2)) for q in cs]
        return B.Leg(tuple(cs), tuple(dims), 1)

    mpo, _ = BM.to_block_mpo(xxz_mpo(2 * nh + 2), [[1, -1]] * (2 * nh + 2), dev)
    Wf, Wb, Wl = mpo[0], mpo[1], mpo[-1]
    p = Wb.legs[2]
    rng = np.random.default_rng(seed)
    legs = [leg(n) for n in range(nh + 1)]
    L = BO.BlockTensor(B.canonical((legs[0], Wf.legs[0].flip(), legs[0].flip()), 0),
                       torch.ones(1, dtype=torch.float64, device=dev))
    for n in range(nh):
        A = BO.random((legs[n], p, legs[n + 1].flip()), 0, rng, dev)
        L = BM.update_left_env(L, A, Wf if n == 0 else Wb)
        L = L.with_data(L.data / L.data.abs().max())
    R = BO.BlockTensor(B.canonical((legs[0].flip(), Wl.legs[1].flip(), legs[0]), 0),
                       torch.ones(1, dtype=torch.float64, device=dev))
    for n in range(nh):
        Bt = BO.random((legs[n + 1], p, legs[n].flip()), 0, rng, dev)
        R = BM.update_right_env(R, Bt, Wl if n == 0 else Wb)
        R = R.with_data(R.data / R.data.abs().max())
    L, R = L.canonical(), R.canonical()
    th = BO.random((legs[nh], p, p, legs[nh].flip()), 0, rng, dev)
    th = th.with_data(th.data / th.data.norm())
    return dict(L=L, R=R, W=Wb, theta=th, leg=legs[nh], sectors=target)


def block_bench(steps, warmup, chi=2048, local=True):
    """Config of the U(1) XXZ chain at chi = 2048 over 31 Gaussian sectors: the block H_eff
    matvec in float64 and with the float32 operator (float64 vectors), per-kernel events, hipGraph
    replay, and the local ground state by float64 vs mixed Lanczos."""
    import torch
    from tnkit_amd import block_ops as BO
    from tnkit_amd.eigensolver import lanczos_ground_mixed, lanczos_ground_restart
    t0 = time.perf_counter()
    pr = build_block_problem(chi)
    L, R, W, th = pr["L"], pr["R"], pr["W"], pr["theta"]
    op64 = BO.BlockTwoSite(L, W, W, R)
    op32 = BO.BlockTwoSite(L, W, W, R, dtype=torch.float32)
    op64.tune()
    op32.tune()
    setup = time.perf_counter() - t0
    x = th.data
    out = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    y64 = op64.apply(x)
    y32 = op32.apply(x)
    rel32 = float((y32 - y64).abs().max() / y64.abs().max())
    res = {"workload": f"U(1) XXZ (2Sz=0) bulk two-site problem, chi={chi} per bond over "
                       f"{len(pr['sectors'])} sectors (Gaussian, largest {max(pr['sectors'].values())}, "
                       f"smallest {min(pr['sectors'].values())}), theta {op64.size} elements",
           "gflop_per_matvec": op64.flops / 1e9, "macs_per_contraction": op64.macs,
           "groups_per_contraction": [len(st[0].desc) for st in op64.stages],
           "tile_cfg_f64": [st[4] for st in op64.stages],
           "tile_cfg_f32": [st[4] for st in op32.stages], "setup_s": setup,
           "matvec_rel_err_f32": rel32}
    for name, op in (("fp64", op64), ("mixed", op32)):
        for _ in range(warmup):
            op.apply(x, out)
        ms = event_time(lambda: op.apply(x, out), steps, stream)
        run = op.graph()
        for _ in range(warmup):
            run(x)
        msg = event_time(lambda: run(x), steps, stream)
        stages = []
        for i, (p, A, Bb, o, cfg) in enumerate(op.stages):
            src = x if i == 0 else op.stages[i - 1][3]
            tg = event_time(lambda: BO.gather(p.dev_b, src, op.dtype, Bb), steps, stream)
            tm = event_time(lambda: BO.grouped_gemm(p, A, Bb, o, cfg), steps, stream)
            esz = 4 if op.dtype == torch.float32 else 8
            mov = (src.numel() * src.element_size() + Bb.numel() * esz) / (tg * 1e-3) / 1e9
            stages.append({"gather_ms": round(tg, 4), "gather_GBps": round(mov, 1),
                           "gemm_ms": round(tm, 4),
                           "gemm_tflops": round(2 * p.macs / (tm * 1e-3) / 1e12, 2),
                           "cfg": cfg})
        src = op.stages[-1][3]
        tf = event_time(lambda: BO.gather(op.final, src, op.io_dtype, out), steps, stream)
        res[name] = {"ms": ms, "tflops": op.flops / (ms * 1e-3) / 1e12, "graph_ms": msg,
                     "graph_tflops": op.flops / (msg * 1e-3) / 1e12, "stages": stages,
                     "final_gather_ms": round(tf, 4)}
    res["speedup_mixed_vs_fp64"] = res["fp64"]["ms"] / res["mixed"]["ms"]
    if local:
        cnt = {"lo": 0, "hi": 0}
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        e64, _, i64 = lanczos_ground_restart(op64.apply, x.clone(), tol=1e-9, max_iter=4 * chi)
        torch.cuda.synchronize()
        t_e64 = time.perf_counter() - t1
        t1 = time.perf_counter()
        emx, _, _ = lanczos_ground_mixed(op32.apply, op64.apply, x.clone(), tol=1e-9,
                                         max_iter=4 * chi, counts=cnt)
        torch.cuda.synchronize()
        res["local_dE"] = {"fp64_E": e64, "mixed_E": emx, "abs": abs(emx - e64),
                           "fp64_matvecs": i64["iters"], "mixed_lo": cnt["lo"],
                           "mixed_hi": cnt["hi"], "fp64_s": t_e64,
                           "mixed_s": time.perf_counter() - t1}
    return res


def sweep_cost(N, chis, forms, log=None):
    """tnkit's per-sweep cost study (BENCHMARKS.md, "Cost per DMRG sweep"): the Heisenberg chain
    (XXZ, Jxy = Jz = 1) of N sites, exactly 3 sweeps from a random state at full chi, seconds per
    sweep as the mean of sweeps 2-3; dense and U(1) block forms (2 S^z_total = 0)."""
    import torch
    from tnkit_amd import _dmrg as H
    from tnkit_amd import block_mps as BM
    from tnkit_amd.models import xxz_mpo
    dense_mpo = xxz_mpo(N)
    dev = torch.device("cuda", 0)
    block_mpo, _ = BM.to_block_mpo(dense_mpo, [[1, -1]] * N, dev)
    out = []
    for form in forms:
        for chi in chis:
            kw = dict(n_sweeps=3, tol=0.0, rtol=0.0, seed=0, expand=0)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if form.startswith("u1"):
                E, _, info = H.dmrg(block_mpo, chi, charge=0, eig=form.split(":")[1],
                                    trunc="gram", **kw)
            else:
                E, _, info = H.dmrg(dense_mpo, chi, backend="hip-real", trunc="gram",
                                    eig=form.split(":")[1], **kw)
            rec = {"form": form, "chi": chi, "E": E, "sweep_s": info["sweep_s"],
                   "s_per_sweep_2_3": sum(info["sweep_s"][1:3]) / 2,
                   "wall_s": time.perf_counter() - t0, "matvec_calls": info["matvec_calls"],
                   "phase_s": {k: round(v, 2) for k, v in info["phase_time"].items()},
                   "rel_err_vs_ref": (E + 28.17542485974319) / 28.17542485974319}
            out.append(rec)
            if log:
                print(json.dumps(rec), file=log, flush=True)
    return out


def small_linalg(reps=20):
    """The small factorisations of the truncations on the device: the Jacobi kernels
    (kernels.eigh_jacobi, kernels.svd_jacobi) against torch.linalg (rocSOLVER) on the shapes the
    sweeps meet: a batch of 32 complex Gram matrices of order 50 (the batched TDVP at bonds 25)
    and float64 SVDs of 16 x 64, 32 x 32 and 64 x 64 (two-site matrices at bonds 8-32); the
    library's float64 eigensolver at orders 512..2048; its batched complex eigensolver and QR on
    the batched TDVP's truncation shapes at binding caps. Milliseconds per call with HIP events;
    every result is checked against torch's."""
    import torch
    from tnkit_amd import kernels
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    out = {"eigh": [], "svd": []}
    X = torch.randn((32, 50, 50), dtype=torch.complex128, device=dev, generator=g)
    G = (X @ X.mH.resolve_conj()).contiguous()
    w, _, sw = kernels.eigh_jacobi(G)
    err = float((w - torch.linalg.eigvalsh(G)).abs().max() / w.abs().max())
    out["eigh"].append({"shape": [32, 50, 50], "dtype": "c128",
                        "jacobi_ms": event_time(lambda: kernels.eigh_jacobi(G), reps, stream),
                        "torch_ms": event_time(lambda: torch.linalg.eigh(G), reps, stream),
                        "max_sweeps": int(sw.max()), "rel_err_eigenvalues": err})
    for m, n in ((16, 64), (32, 32), (64, 64)):
        A = torch.randn((m, n), dtype=torch.float64, device=dev, generator=g)
        _, S, _, sw = kernels.svd_jacobi(A)
        ref = torch.linalg.svdvals(A)
        out["svd"].append({"shape": [m, n], "dtype": "f64",
                           "jacobi_ms": event_time(lambda: kernels.svd_jacobi(A), reps, stream),
                           "torch_ms": event_time(lambda: torch.linalg.svd(A, full_matrices=False),
                                                  reps, stream),
                           "sweeps": int(sw.max()),
                           "rel_err_singular_values": float((S - ref).abs().max() / ref[0])})
    # the library's eigensolver for the Gram truncation at orders 65..2048 (all n eigenpairs,
    # float64 Gram matrices with decaying columns) against torch.linalg.eigh (rocSOLVER syevd)
    out["syev"] = []
    for n in (512, 1024, 2048):
        X = torch.randn((n, n), dtype=torch.float64, device=dev, generator=g)
        X = X * torch.exp(-torch.arange(n, dtype=torch.float64, device=dev) / (n / 20.0))
        A = (X @ X.T).contiguous()
        A = 0.5 * (A + A.T)
        w, Z = kernels.syev_top(A, n)
        ref = torch.linalg.eigvalsh(A)
        lam = torch.flip(w, dims=[0])
        out["syev"].append({
            "n": n, "k": n, "dtype": "f64",
            "own_ms": event_time(lambda: kernels.syev_top(A, n, check_status=False), reps, stream),
            "torch_ms": event_time(lambda: torch.linalg.eigh(A), reps, stream),
            "rel_err_eigenvalues": float((w - ref).abs().max() / ref.abs().max()),
            "rel_residual": float((A @ Z - Z * lam[None, :]).abs().max() / ref.abs().max()),
            "orthogonality": float((Z.T @ Z - torch.eye(n, dtype=torch.float64, device=dev))
                                   .abs().max())})
    # the batched TDVP's truncation at binding caps: 32 complex Gram matrices of order 256 and
    # 512 (all eigenpairs) and 32 complex QRs of the projected factor (512 x 128, 512 x 256)
    out["heev"], out["qr"] = [], []
    for n in (256, 512):
        X = torch.randn((32, n, n), dtype=torch.complex128, device=dev, generator=g)
        X = X * torch.exp(-torch.arange(n, dtype=torch.float64, device=dev) / (n / 20.0))
        A = (X @ X.mH.resolve_conj()).contiguous()
        A = 0.5 * (A + A.mH)
        w, Z = kernels.heev_top(A, n)
        ref = torch.linalg.eigvalsh(A)
        lam = torch.flip(w, dims=[1])
        sc = ref.abs().amax(dim=1, keepdim=True)
        out["heev"].append({
            "batch": 32, "n": n, "k": n, "dtype": "c128",
            "own_ms": event_time(lambda: kernels.heev_top_async(A, n), reps, stream),
            "torch_ms": event_time(lambda: torch.linalg.eigh(A), reps, stream),
            "rel_err_eigenvalues": float(((w - ref).abs() / sc).max()),
            "rel_residual": float(((A @ Z - Z * lam[:, None, :]).abs().amax(dim=(1, 2))
                                   / sc[:, 0]).max()),
            "orthogonality": float((Z.mH @ Z - torch.eye(n, dtype=Z.dtype, device=dev))
                                   .abs().max())})
    for m, n in ((512, 128), (512, 256)):
        A = torch.randn((32, m, n), dtype=torch.complex128, device=dev, generator=g)
        Q, R = kernels.qr(A)
        Qr, Rr = torch.linalg.qr(A)
        out["qr"].append({
            "batch": 32, "m": m, "n": n, "dtype": "c128",
            "own_ms": event_time(lambda: kernels.qr(A, check_status=False), reps, stream),
            "torch_ms": event_time(lambda: torch.linalg.qr(A), reps, stream),
            "max_diff_R_vs_torch": float((R - Rr).abs().max()),
            "max_diff_Q_vs_torch": float((Q - Qr).abs().max()),
            "orthogonality": float((Q.mH @ Q - torch.eye(n, dtype=Q.dtype, device=dev))
                                   .abs().max())})
    return out


def config4(legs=32, N=64, n_steps=40, dt=0.05, meas_every=10, svd_tol=1e-12, log=None,
            forms=("c128", "c64"), ragged="pad"):
    """Batched time evolution: `legs` TFIM quenches (N sites) from DMRG ground states at chi = 16
    for fields g0 ~ U[1.2, 2.0] (default_rng(1)) to g = 1, two-site TDVP with per-leg bond caps
    spread over 64..256 and a relative discarded-weight budget svd_tol per cut (so each leg's
    bond follows its entanglement up to its cap; svd_tol = 0 keeps every leg at its cap once the
    bond reaches it), complex128 and complex64-operator (complex128 Krylov) forms (`forms`), the
    batch held as `ragged` ('pad', 'tight' or 'group'), <Z_j>(t) against the exact free-fermion
    curves and, in the complex128 form, against single-leg tdvp2_evolve for two legs."""
    import numpy as np
    import torch
    from tnkit_amd import _dmrg as H
    from tnkit_amd import free_fermion as F
    from tnkit_amd import tdvp as T
    from tnkit_amd.models import tfim_mpo
    from tnkit_amd.observables import local_expectations
    Z = np.diag([1.0, -1.0])
    g0 = np.random.default_rng(1).uniform(1.2, 2.0, legs)
    caps = [int(c) for c in np.linspace(64, 256, legs).round()]
    t0 = time.perf_counter()
    states = []
    for j, g in enumerate(g0):
        _, mps, _ = H.dmrg(tfim_mpo(N, float(g)), 16, n_sweeps=4, tol=1e-12, rtol=0.0, seed=j,
                           backend="hip-real")
        states.append([t.to(torch.complex128) for t in mps])
        if log:
            print(f"ground state {j} g0={g:.3f}", file=log, flush=True)
    t_gs = time.perf_counter() - t0
    times = [k * meas_every * dt for k in range(n_steps // meas_every + 1)]
    ff = [F.quench_z(N, float(g), 1.0, times) for g in g0]
    mpo1 = tfim_mpo(N, 1.0)
    meas = lambda ss: [local_expectations(s, Z) for s in ss]  # noqa: E731
    out = {"workload": f"{legs} TFIM quenches N={N}, g0~U[1.2,2.0] (seed 1) -> g=1, TDVP2 "
                       f"dt={dt} x {n_steps} steps, per-leg caps 64..256, svd_tol={svd_tol}, "
                       f"ragged={ragged}",
           "legs": legs, "N": N, "caps": caps, "ground_state_s": t_gs}
    seq = {}
    runs = [(nm, low) for nm, low in (("c128", False), ("c64_krylov_c128", True))
            if nm.split("_")[0] in forms]
    for name, low in runs:
        stats = {}
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        prog = (lambda k, t: print(f"{name} step {k} t={t:.2f} "
                                   f"{time.perf_counter() - t1:.1f}s", file=log, flush=True)) \
            if log else None
        res, rec = T.evolve_batch(states, mpo1, dt, n_steps, chi_max=caps, krylov_m=20,
                                  keep_device=True, measure=meas, meas_every=meas_every,
                                  stats=stats, low_precision=low, progress=prog,
                                  svd_tol=svd_tol, ragged=ragged)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        err = max(float(np.abs(np.asarray(zs[j]) - ff[j][k]).max())
                  for k, (_, zs) in enumerate(rec) for j in range(legs))
        out[name] = {"wall_s": wall, "s_per_step": wall / n_steps,
                     "leg_steps_per_s": legs * n_steps / wall,
                     "z_max_err_vs_free_fermion": err, "krylov_m": stats.get("krylov_m"),
                     "max_bond": max(max(b) for b in stats["bonds"]),
                     "z_final_leg0": [float(v) for v in rec[-1][1][0][:4]]}
        if not low:
            for j in (0, legs - 1):
                _, r1 = T.tdvp2_evolve(states[j], mpo1, caps[j], dt, n_steps, krylov_m=20,
                                       measure=lambda m: local_expectations(m, Z),
                                       meas_every=meas_every, svd_tol=svd_tol)
                seq[j] = r1
            out[name]["z_max_diff_vs_sequential"] = max(
                float(np.abs(np.asarray(rec[k][1][j]) - np.asarray(seq[j][k][1])).max())
                for j in seq for k in range(len(rec)))
        if log:
            print(json.dumps({name: out[name]}), file=log, flush=True)
    return out


def config4_modes(legs=32, N=64, n_steps=4, dt=0.05, log=None):
    """The batched TDVP's ragged modes on binding caps: `legs` TFIM quenches as in `config4`
    with svd_tol = 0 (every leg keeps its cap once its bond reaches it), run for n_steps in
    'pad' (fixed widest shapes), 'tight' (widest kept rank) and 'group' (one batch per entry
    bond profile). Leg-steps per second per mode, and the largest difference of the final
    <Z_j> from the 'pad' run."""
    import numpy as np
    import torch
    from tnkit_amd import _dmrg as H
    from tnkit_amd import tdvp as T
    from tnkit_amd.models import tfim_mpo
    from tnkit_amd.observables import local_expectations
    Z = np.diag([1.0, -1.0])
    g0 = np.random.default_rng(1).uniform(1.2, 2.0, legs)
    caps = [int(c) for c in np.linspace(64, 256, legs).round()]
    states = []
    for j, g in enumerate(g0):
        _, mps, _ = H.dmrg(tfim_mpo(N, float(g)), 16, n_sweeps=4, tol=1e-12, rtol=0.0, seed=j,
                           backend="hip-real")
        states.append([t.to(torch.complex128) for t in mps])
        if log:
            print(f"ground state {j} g0={g:.3f}", file=log, flush=True)
    mpo1 = tfim_mpo(N, 1.0)
    meas = lambda ss: [local_expectations(s, Z) for s in ss]  # noqa: E731
    out = {"workload": f"{legs} TFIM quenches N={N} -> g=1, TDVP2 dt={dt} x {n_steps} steps, "
                       "per-leg caps 64..256, svd_tol=0", "caps": caps}
    zref = None
    for mode in ("pad", "tight", "group"):
        stats = {}
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        prog = (lambda k, t, mode=mode, t1=t1: print(
            f"{mode} step {k} t={t:.2f} {time.perf_counter() - t1:.1f}s", file=log,
            flush=True)) if log else None
        _, rec = T.evolve_batch(states, mpo1, dt, n_steps, chi_max=caps, krylov_m=20,
                                keep_device=True, measure=meas, meas_every=n_steps,
                                stats=stats, svd_tol=0.0, ragged=mode, progress=prog)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        z = np.asarray(rec[-1][1])
        if zref is None:
            zref = z
        out[mode] = {"wall_s": wall, "leg_steps_per_s": legs * n_steps / wall,
                     "final_bonds_max": max(max(b) for b in stats["bonds"]),
                     "z_max_diff_vs_pad": float(np.abs(z - zref).max())}
        if log:
            print(json.dumps({mode: out[mode]}), file=log, flush=True)
    out["tight_vs_pad_speedup"] = out["tight"]["leg_steps_per_s"] / out["pad"]["leg_steps_per_s"]
    out["group_vs_pad_speedup"] = out["group"]["leg_steps_per_s"] / out["pad"]["leg_steps_per_s"]
    return out


def config5(N=64, chi=256, fd_sites=(0, 21, 42, 63), h=1e-5, log=None):
    """Differentiation: the TFIM with per-site fields g_i = 1 + 0.1 N(0,1) (default_rng(2)),
    two-site DMRG at chi (float64), then dE/dg_i by the hand-written backward kernels
    (MPOExpectation), compared with -<Z_i> (Hellmann-Feynman), with central differences of the
    exact free-fermion ground energy, and with central differences of warm-started DMRG energies
    at a few sites."""
    import numpy as np
    import torch
    from tnkit_amd import _dmrg as H
    from tnkit_amd import free_fermion as F
    from tnkit_amd.autodiff import field_gradient
    from tnkit_amd.models import tfim_mpo_fields
    from tnkit_amd.observables import local_expectations
    g = 1.0 + 0.1 * np.random.default_rng(2).standard_normal(N)
    opts = dict(n_sweeps=6, tol=1e-12, rtol=0.0, seed=0, backend="hip-real", trunc="gram")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, mps, info = H.dmrg(tfim_mpo_fields(g), chi, **opts)
    torch.cuda.synchronize()
    t_dmrg = time.perf_counter() - t0
    if log:
        print(f"dmrg E={E:.14f} sweeps={info['sweeps']} {t_dmrg:.1f}s", file=log, flush=True)
    t0 = time.perf_counter()
    grad = field_gradient(list(mps), torch.as_tensor(g)).cpu().numpy()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    z = local_expectations(list(mps), np.diag([1.0, -1.0]))
    E_exact = F.tfim_ground_energy(N, g)
    fd_exact = np.empty(N)
    for i in range(N):
        e = np.zeros(N)
        e[i] = h
        fd_exact[i] = (F.tfim_ground_energy(N, g + e) - F.tfim_ground_energy(N, g - e)) / (2 * h)
    fd_dmrg = {}
    for i in fd_sites:
        e = np.zeros(N)
        e[i] = h
        Ep, _, _ = H.dmrg(tfim_mpo_fields(g + e), chi, **dict(opts, n_sweeps=3,
                                                                init_mps=list(mps)))
        Em, _, _ = H.dmrg(tfim_mpo_fields(g - e), chi, **dict(opts, n_sweeps=3,
                                                                init_mps=list(mps)))
        fd_dmrg[i] = (Ep - Em) / (2 * h)
        if log:
            print(f"fd site {i}: {fd_dmrg[i]:.10f} vs grad {grad[i]:.10f}", file=log, flush=True)
    return {"workload": f"TFIM N={N}, g_i = 1 + 0.1 N(0,1) (seed 2), two-site DMRG chi={chi}",
            "E": E, "E_exact_free_fermion": E_exact, "dE_abs": abs(E - E_exact),
            "dmrg_s": t_dmrg, "gradient_s": t_grad,
            "max_abs_grad_plus_z": float(np.abs(grad + z).max()),
            "max_abs_grad_minus_exact_fd": float(np.abs(grad - fd_exact).max()),
            "max_abs_grad_minus_dmrg_fd": float(max(abs(grad[i] - v) for i, v in
                                                    fd_dmrg.items())),
            "fd_h": h, "fd_sites": list(fd_sites)}


def config2(N, stages, eig, trunc, seed=0, log=None):
    """The precision study: XXZ chain N sites, dense two-site DMRG with the bond ramped through
    `stages` = [(chi, sweeps), ...], each stage started from the previous state (init_mps), as
    tnkit's benchmark runner chains its stages (bench/suite/runners/tnkit_runner.py:143-152).
    Returns per-stage energies, sweep times, matvec counts and phase times."""
    import torch
    from tnkit_amd import _dmrg as H
    from tnkit_amd.models import xxz_mpo
    mpo = xxz_mpo(N)
    out, mps = [], None
    for chi, n in stages:
        kw = dict(n_sweeps=n, tol=0.0, rtol=0.0, seed=seed, backend="hip-real", eig=eig,
                  trunc=trunc)
        if mps is not None:
            kw["init_mps"] = list(mps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E, mps, info = H.dmrg(mpo, chi, **kw)
        torch.cuda.synchronize()
        later = info["sweep_s"][1:] or info["sweep_s"]
        rec = {"chi": chi, "sweeps": n, "E": E, "energies": info["energies"],
               "sweep_s": info["sweep_s"], "s_per_sweep_after_first": sum(later) / len(later),
               "wall_s": time.perf_counter() - t0,
               "matvec_calls": info["matvec_calls"], "mixed_counts": info["mixed_counts"],
               "phase_s": {k: round(v, 3) for k, v in info["phase_time"].items()},
               "discarded_weight": info["discarded_weight"],
               "max_bond": max(int(t.shape[2]) for t in mps)}
        out.append(rec)
        if log:
            print(json.dumps({"eig": eig, **rec}), file=log, flush=True)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--chi", type=int, default=1024)
    ap.add_argument("--N", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-local", action="store_true",
                    help="skip the local ground-state solves (short profiled runs)")
    ap.add_argument("--config2", type=str, default="",
                    help="run the chi-ramp precision study instead, e.g. '64:4,512:4,1024:4'")
    ap.add_argument("--trunc", type=str, default="gram")
    ap.add_argument("--sweep-cost", type=str, default="",
                    help="per-sweep cost study, e.g. '256,512,1024' (Heisenberg chain, N=--N)")
    ap.add_argument("--forms", type=str, default="u1:lanczos-restart,dense:lanczos-restart")
    ap.add_argument("--config4", action="store_true",
                    help="batched TDVP study: 32 TFIM quenches N=64, per-leg caps 64..256")
    ap.add_argument("--legs", type=int, default=32)
    ap.add_argument("--config4-modes", action="store_true",
                    help="batched TDVP ragged modes pad/tight/group on binding caps (svd_tol=0)")
    ap.add_argument("--config5", action="store_true",
                    help="differentiation study: TFIM N=64 per-site fields, chi=256")
    ap.add_argument("--tsteps", type=int, default=40)
    ap.add_argument("--svd-tol", type=float, default=1e-12,
                    help="--config4: relative discarded-weight budget per cut (0: caps bind)")
    ap.add_argument("--c4-forms", type=str, default="c128,c64",
                    help="--config4: forms to run, from c128 (complex128) and c64 (complex64 "
                         "operator, complex128 Krylov)")
    ap.add_argument("--ragged", type=str, default="pad",
                    help="--config4: how the batch holds uneven bonds: pad, tight or group")
    ap.add_argument("--block", action="store_true",
                    help="only the U(1) block-sparse matvec at chi=2048 (grouped launches)")
    ap.add_argument("--no-block", action="store_true")
    ap.add_argument("--no-linalg", action="store_true",
                    help="skip the truncation factorisations against rocSOLVER (small_linalg)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the chi=1024 sweep-time measurement (Heisenberg N=64, 3 sweeps)")
    ap.add_argument("--small-linalg", action="store_true",
                    help="only time the Jacobi eigh / thin SVD kernels and the order-512..2048 eigensolver "
                         "against torch (rocSOLVER)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="only run a device copy of MB megabytes (counter calibration)")
    args = ap.parse_args()
    if args.config2:
        stages = [tuple(int(v) for v in st.split(":")) for st in args.config2.split(",")]
        res = {"workload": f"XXZ N={args.N} two-site DMRG, stages {stages}, trunc={args.trunc}"}
        for eig in ("lanczos-restart", "lanczos-mixed"):
            res[eig] = config2(args.N, stages, eig, args.trunc, log=sys.stderr)
        res["dE_per_stage"] = [abs(a["E"] - b["E"]) for a, b in
                               zip(res["lanczos-restart"], res["lanczos-mixed"])]
        print(json.dumps(res))
        return
    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe * 2 ** 20 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": src.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the two-site path is measured on "
                                   "one MI355X and multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    if args.block:
        print(json.dumps(block_bench(args.steps, args.warmup, local=not args.no_local)))
        return
    if args.small_linalg:
        print(json.dumps(small_linalg(max(5, args.steps))))
        return
    if args.config5:
        print(json.dumps(config5(64 if args.N == 100 else args.N, log=sys.stderr)))
        return
    if args.config4_modes:
        print(json.dumps(config4_modes(args.legs, 64 if args.N == 100 else args.N,
                                       4 if args.tsteps == 40 else args.tsteps, log=sys.stderr)))
        return
    if args.config4:
        print(json.dumps(config4(args.legs, 64 if args.N == 100 else args.N, args.tsteps,
                                 svd_tol=args.svd_tol, log=sys.stderr,
                                 forms=tuple(args.c4_forms.split(",")), ragged=args.ragged)))
        return
    if args.sweep_cost:
        chis = [int(c) for c in args.sweep_cost.split(",")]
        res = {"workload": f"Heisenberg chain N={args.N}, 3 sweeps from random at full chi, "
                           "s/sweep = mean of sweeps 2-3",
               "runs": sweep_cost(args.N, chis, args.forms.split(","), log=sys.stderr)}
        print(json.dumps(res))
        return

    cores = len(os.sched_getaffinity(0))
    for v in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ.setdefault(v, str(cores))
    import numpy as np
    import torch
    from tnkit_amd import kernels
    from tnkit_amd import _dmrg as H

    t_setup = time.perf_counter()
    prob = build_local_problem(args.N, args.chi)
    L, R, th = prob["L"], prob["R"], prob["theta"]
    W = prob["mpo"][prob["i0"]]
    a, d1, d2, b = th.shape
    wl, wr = W.shape[0], W.shape[1]
    fl = two_site_flops(a, b, d1, d2, wl, wr)
    stream = torch.cuda.current_stream()
    op64 = kernels.fused_two_site(L, W, W, R)
    op32 = kernels.fused_two_site(L, W, W, R, dtype=torch.float32)
    op16 = kernels.fused_two_site(L, W, W, R, dtype=torch.bfloat16)
    ws = kernels.two_site_ws(th, *op64)
    out = torch.empty_like(th)

    def mv32():
        kernels.apply_two_site_fused(th, *op32, out=out, ws=ws)

    def mv64():
        kernels.apply_two_site_fused(th, *op64, out=out, ws=ws)

    def mv16():
        kernels.apply_two_site_fused(th, *op16, out=out, ws=ws)

    # accuracy of one mixed matvec against float64
    y64 = kernels.apply_two_site_fused(th, *op64)
    y32 = kernels.apply_two_site_fused(th, *op32)
    mv_rel = float((y32 - y64).abs().max() / y64.abs().max())
    y16 = kernels.apply_two_site_fused(th, *op16)
    mv_rel16 = float((y16 - y64).abs().max() / y64.abs().max())
    del y16

    # ---- timed region: K mixed matvecs --------------------------------------------------
    for _ in range(args.warmup):
        mv32()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms32_ev = event_time(mv32, args.steps, stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms32 = wall * 1e3 / args.steps
    for _ in range(args.warmup):
        mv64()
    ms64 = event_time(mv64, args.steps, stream)
    for _ in range(args.warmup):
        mv16()
    ms16 = event_time(mv16, args.steps, stream)
    st16 = {}
    for st in (1, 2, 4):
        fn = lambda st=st: kernels.apply_two_site_fused(th, *op16, out=out, ws=ws, stages=st)  # noqa
        fn()
        st16[st] = event_time(fn, max(5, args.steps // 2), stream)

    # ---- per-kernel events for the roofline ----------------------------------------------
    sflops = stage_flops(a, b, d1, d2, wl, wr)
    sbytes = stage_bytes(a, b, d1, d2, wl, wr, 4, 8)
    # the mixed matvec runs stages 1 + 2 as one kernel (GEMM1 with the MPO mix in its
    # epilogue) and stage 4; stages 1 and 2 alone are the unfused kernels, timed for reference
    stage_ms = {}
    for st in (1, 2, 3, 4):
        fn = lambda st=st: kernels.apply_two_site_fused(th, *op32, out=out, ws=ws, stages=st)  # noqa
        fn()
        stage_ms[st] = event_time(fn, max(5, args.steps // 2), stream)
    # the fused kernel alone: theta already in float32 (the mixed matvec narrows it first, in
    # narrow_kernel, which stage 3 above includes)
    th32 = th.to(torch.float32).contiguous()
    out32 = torch.empty_like(th32)
    fn = lambda: kernels.apply_two_site_fused(th32, *op32, out=out32, ws=ws, stages=3)  # noqa
    fn()
    stage_ms[5] = event_time(fn, max(5, args.steps // 2), stream)
    del th32, out32
    names = {1: "GEMM1 theta x L, unfused (narrow_kernel + gemm_tn_kernel<float,float,float,float>)",
             2: "MPO mix, unfused (mix_mfma_kernel<float,float,float>)",
             3: "narrow + GEMM1+mix (narrow_kernel + gemm1_mix_dma_kernel<float,5,2>)",
             4: "GEMM3 x R (gemm_dma_kernel<float,double,32,2>)",
             5: "GEMM1+mix theta x L with the MPO mix in the epilogue "
                "(gemm1_mix_dma_kernel<float,5,2>)"}
    sflops[5] = sflops[3]
    dom = max((5, 4), key=stage_ms.get)
    if dom == 2:
        achieved = sbytes[2] / (stage_ms[2] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK, "unit": "GB/s",
                "frac": achieved / HBM_PEAK, "traffic": None}
    else:
        achieved = sflops[dom] / (stage_ms[dom] * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": PEAK["f32"], "unit": "TFLOP/s",
                "frac": achieved / PEAK["f32"], "traffic": None}
    roof["kernel"] = names[dom]
    roof["traffic"], roof["traffic_source"] = measured_traffic(names[dom], a, "mixed")
    roof["per_stage_ms"] = {names[k]: round(v, 4) for k, v in stage_ms.items()}

    # ---- local ground state: mixed vs float64 -------------------------------------------
    local = {}
    if not args.no_local:
        local = local_solves(th, op64, op32, ws, args.chi, op16=op16)
    del op64, op32, op16, ws, out, y32, y64
    cgs2 = cgs2_bench(th.numel())

    # the truncation's factorisations on the library's solvers against rocSOLVER (Jacobi eigh
    # and SVD, the order 512..2048 eigensolver, the batched complex eigensolver and QR)
    linalg_res = {} if args.no_linalg else small_linalg(5)
    setup_s = time.perf_counter() - t_setup
    block = {} if args.no_block else block_bench(max(5, args.steps // 2), args.warmup,
                                                 local=not args.no_local)
    # the sweep half of the metric: config 2's bond ramp on the XXZ chain N = 100, measured in
    # both precisions (each stage started from the previous state, as tnkit's runner chains them)
    ramp = {}
    if not args.no_sweep:
        stages = [(64, 4), (512, 4), (args.chi, 4)]
        ramp = {"workload": f"XXZ chain N={args.N}, two-site DMRG, bond ramp {stages} "
                            "(12 sweeps), trunc=gram, seed 0",
                "reference_h100_mig_fp64_s_per_sweep_chi1024_N64": 38.4}
        for eig in ("lanczos-restart", "lanczos-mixed"):
            ramp[eig] = config2(args.N, stages, eig, "gram", log=sys.stderr)
        ramp["dE_per_stage"] = [abs(x["E"] - y["E"]) for x, y in
                                zip(ramp["lanczos-restart"], ramp["lanczos-mixed"])]
        ramp["s_per_sweep_at_chi"] = {
            eig: ramp[eig][-1]["s_per_sweep_after_first"] for eig in
            ("lanczos-restart", "lanczos-mixed")}
        ramp["sweep_speedup_mixed_vs_fp64"] = (ramp["s_per_sweep_at_chi"]["lanczos-restart"]
                                               / ramp["s_per_sweep_at_chi"]["lanczos-mixed"])

    res = {
        "metric": METRIC, "value": fl / (ms32 * 1e-3) / 1e12, "unit": "TFLOP/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms32,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"XXZ chain N={args.N} (Jxy=Jz=1, D=5, d=2): H_eff matvec on "
                               f"the bulk two-site problem at chi={args.chi}, environments of a "
                               f"random mixed-canonical MPS; mixed = float32 MFMA operator with "
                               f"float64 vectors", "chi": args.chi, "N": args.N, "D": 5, "d": 2},
        "gflop_per_step": fl / 1e9,
        "event_ms_per_step": ms32_ev,
        "fp64": {"ms": ms64, "tflops": fl / (ms64 * 1e-3) / 1e12},
        "speedup_mixed_vs_fp64": ms64 / ms32,
        "matvec_rel_err_f32": mv_rel,
        "bf16": {"ms": ms16, "tflops": fl / (ms16 * 1e-3) / 1e12, "matvec_rel_err": mv_rel16,
                 "speedup_vs_mixed": ms32 / ms16,
                 "per_stage_ms": {"GEMM1 (gemm_tn_bf16_kernel<double,bf16,float>)": st16[1],
                                  "MPO mix (float, bf16 out)": st16[2],
                                  "GEMM3 (gemm_tn_bf16_kernel<bf16,bf16,double>)": st16[4]},
                 "note": "explored, not the headline: bf16 operator (LT, RT) and bf16 theta / "
                         "mixed intermediate on the bf16 matrix cores, float accumulation"},
        "local_dE": local,
        "reference_h100_mig_fp64_tflops": 5.67,
        "roofline": roof,
        "cgs2": cgs2,
        "block_u1": block,
        "config2_ramp": ramp,
        "truncation_linalg": linalg_res,
        "setup_s": setup_s,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(prob)
        res["config1"] = config1_cpu_vs_gpu()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
