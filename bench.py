"""Benchmark of the iterative-solve path on one MI355X.

Workload: the SPE10-shaped 60x220x85 system (1,122,000 cells, 7,780,000
nonzeros) assembled with the reference's own formula (modules/mesh/spe.cc:69-131,
reaction coefficient c = 1) on a seeded synthetic permeability field, solved by
PCG with the default uh_cheb preconditioner (sigma 3, 2 Chebyshev steps) from
x0 = 0 to a relative residual of 1e-8.  A step is one complete solve with the
matrix, hierarchy and right-hand side already resident in HBM.

Prints one JSON line: the mixed-precision solve time (FP64 outer flexible PCG,
FP32 hierarchy) as `value`, the FP64 solve time beside it, the SpMV roofline
(cold caches) and the single-core CPU oracle timed on the same host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "PCG time to 1e-8 rel-res @1.12M cells, mixed vs FP64 GPU; SpMV HBM GB/s"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """The host's CPU model and logical CPU count (recorded beside every CPU baseline)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{model}, {os.cpu_count()} logical CPUs visible"


def load_profile(name):
    """A committed rocprofv3 PMC summary (profiles/), or None."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def load_traffic():
    """HBM bytes per SpMV launch from the committed rocprofv3 PMC summary, if present."""
    d = load_profile("k_spmv_spe10_60x220x85_traffic.json")
    return d.get("traffic_bytes_per_launch") if d else None


def load_apply_traffic(prec):
    """DRAM bytes per preconditioner apply (FETCH_SIZE + WRITE_SIZE of its level kernels), if measured."""
    d = load_profile("prec_apply_uh_cheb_spe10_60x220x85_traffic.json")
    return d[prec].get("dram_bytes_per_apply") if d and prec in d else None


def cpu_baseline(s, eps, budget_s=20.0):
    """The C oracle (single thread) on the same system: construction + PCG from x0 = 0."""
    import oracle as O
    t0 = time.perf_counter()
    P = O.cheb_prec(s.ia, s.ja, s.a)
    t_const = time.perf_counter() - t0
    times, iters = [], 0
    while True:
        t0 = time.perf_counter()
        x, st = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=eps, x0=np.zeros(s.n), maxit=10000)
        times.append(time.perf_counter() - t0)
        iters = st["niter"]
        if sum(times) + t_const > budget_s / 2 or len(times) >= 5:
            break
    return {"value": 1e3 * float(np.median(times)), "unit": "ms per solve", "cores": 1, "kind": "port", "host": host_cpu(),
            "sample": f"{len(times)} full PCG solves of the same 60x220x85 system (x0=0, eps={eps:g}, "
                      f"{iters} iterations each) by the C oracle on one host core; construction {t_const:.2f} s",
            "t_const_s": t_const, "niter": iters}, x


def _ensemble_cpu_job(item):
    """One ensemble member on one host core: the oracle's construction, then one timed PCG solve."""
    nz, seed, eps = item
    import oracle as O
    from ml_prec_amd import spe_synth as S
    s = S.ensemble_system(nz, seed)
    P = O.cheb_prec(s.ia, s.ja, s.a)
    t0 = time.perf_counter()
    _, st = O.pcg(s.ia, s.ja, s.a, s.b, P, eps=eps, x0=np.zeros(s.n), maxit=10000)
    return time.perf_counter() - t0, int(st["niter"])


def ensemble_cpu_baseline(spec, eps, procs):
    """The C oracle on every ensemble member, one process per host core (spawned before this process uses the
    GPU): the aggregate rate is the member count over the summed single-core solve time divided by the processes."""
    import multiprocessing as mp
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(procs) as pool:
        res = pool.map(_ensemble_cpu_job, [(nz, seed, eps) for nz, seed in spec], chunksize=1)
    wall = time.perf_counter() - t0
    tsum = sum(t for t, _ in res)
    return {"value": round(len(spec) / (tsum / procs), 2), "unit": "solves/s", "cores": procs, "kind": "port",
            "host": host_cpu(),
            "sample": f"all {len(spec)} members, one oracle PCG solve each (construction excluded), in {procs} "
                      f"processes; summed single-core solve time {tsum:.2f} s, wall time with construction "
                      f"{wall:.1f} s", "iterations": [it for _, it in res]}


def ensemble(N, S, args, emit=True):
    """Ensemble: COUNT SPE10-shaped systems of 60x220xnz cells, nz drawn from 5/10/20/40/85, each with its own
    field; matrices assembled and hierarchies built on the device.  A step solves all of them to eps from
    x0 = 0.  Headline: the systems grouped by hierarchy depth, each group stacked into one block-diagonal
    system with one hierarchy (FP32, FP64 flexible PCG with true residual every 20 iterations, as the
    configuration asks) and solved by mlp_pcg_ensemble, every launch covering the whole group; the groups run
    concurrently, one stream each, driven from one host thread (mlp_pcg_ensemble_batch).  Beside it: the same stacked groups in FP64, the systems as 64 concurrent
    solves on 64 streams (mlp_pcg_batch, FP64), and one after another."""
    from concurrent.futures import ThreadPoolExecutor
    spec = S.ensemble_spec(args.ensemble)
    cpu = None
    if not args.no_cpu:
        procs = max(1, min(16, os.cpu_count() or 1))
        cpu = ensemble_cpu_baseline(spec, args.eps, procs)
        log(f"CPU oracle: {cpu['value']:.1f} solves/s aggregate on {procs} processes ({cpu['sample']})")
    t0 = time.perf_counter()
    with ThreadPoolExecutor(8) as ex:
        fields = list(ex.map(lambda p: S.spe10_field(60, 220, p[0], p[1]).ravel(), spec))
    rhs = [S.rhs(k.size, seed + 7919) for k, (_, seed) in zip(fields, spec)]
    hh = (S.HX, S.HY, S.HZ)
    # one context per system: its matrix and hierarchy (device assembly and construction)
    members = []
    for (nz, _), kx, b in zip(spec, fields, rhs):
        ctx = N.Context(0)
        A = N.DevMatrix.spe_on_device(ctx, ctx.to_device(kx), (nz, 220, 60), kind="spe", c=1.0, kz_ratio=0.1, h=hh)
        H = N.build_uh_cheb_on_device(ctx, A)
        B = N.DevPrec(ctx, A, H, precision=N.MLP_FP64)
        members.append(dict(ctx=ctx, A=A, B=B, depth=H.nlevels, nz=nz, n=kx.size, b=ctx.to_device(b),
                            x=ctx.array(kx.size)))
    cells = sum(m["n"] for m in members)
    groups = {}
    for i, m in enumerate(members):
        groups.setdefault(m["depth"], []).append(i)
    # the stacked groups, each on a context of its own, in FP64 and FP32
    stacks = []
    for depth, idx in sorted(groups.items()):
        ctx = N.Context(0)
        kx = np.concatenate([fields[i] for i in idx])
        A = N.DevMatrix.spe_ensemble_on_device(ctx, ctx.to_device(kx), 60, 220, [spec[i][0] for i in idx],
                                               kind="spe", c=1.0, kz_ratio=0.1, h=hh)
        H = N.build_uh_cheb_on_device(ctx, A)
        assert H.nlevels == depth
        off = np.concatenate([[0], np.cumsum([members[i]["n"] for i in idx])]).astype(np.uint32)
        # each group's launch schedule picked by the autotuner (candidate schedules timed on the group)
        B64 = N.DevPrec.tuned(ctx, A, H, precision=N.MLP_FP64, reps=5)
        B32 = N.DevPrec.tuned(ctx, A, H, precision=N.MLP_FP32, reps=5)
        stacks.append(dict(ctx=ctx, A=A, B64=B64, B32=B32, idx=idx, off=off, depth=depth,
                           b=ctx.to_device(np.concatenate([rhs[i] for i in idx])), x=ctx.array(int(off[-1]))))
    log(f"ensemble: {len(members)} systems, {cells} cells, {len(stacks)} depth groups "
        f"{[(g['depth'], len(g['idx'])) for g in stacks]}, set up in {time.perf_counter() - t0:.1f}s")
    kw = dict(eps=args.eps, maxit=10000, x0_mode=N.X0_ZERO, check_every=8, use_graph=bool(args.graph))
    kw32 = dict(kw, flexible=True, true_res_every=20)

    def stacked(mixed):
        # every group on its own stream, all driven from this thread (mlp_pcg_ensemble_batch)
        jobs = [(g["ctx"], g["A"], g["B32"] if mixed else g["B64"], g["off"], g["b"], g["x"]) for g in stacks]
        rc, res = N.pcg_ensemble_batch(jobs, **(kw32 if mixed else kw))
        if rc:
            raise RuntimeError(f"ensemble group solve failed: {rc}")
        for g in stacks:
            g["ctx"].sync()
        return [st for _, st in res]

    def timed(fn, steps, warmup):
        for _ in range(max(1, warmup)):
            out = fn()
        t = time.perf_counter()
        for _ in range(steps):
            out = fn()
        return 1e3 * (time.perf_counter() - t) / steps, out

    ms32, st32 = timed(lambda: stacked(True), args.steps, args.warmup)
    x32 = [g["x"].get() for g in stacks]
    ms64, st64 = timed(lambda: stacked(False), args.steps, args.warmup)
    x64 = [g["x"].get() for g in stacks]
    it32 = [d["niter"] for g in st32 for d in g]
    worst32 = max(d["true_rel_res"] for g in st32 for d in g)
    log(f"stacked groups, FP32 hierarchy: {ms32:.2f} ms for {len(members)} solves "
        f"({len(members) / ms32 * 1e3:.0f} solves/s), iterations {min(it32)}..{max(it32)}, worst true rel res "
        f"{worst32:.2e}")
    log(f"stacked groups, FP64: {ms64:.2f} ms ({len(members) / ms64 * 1e3:.0f} solves/s)")
    jobs = [(m["ctx"], m["A"], m["B"], m["b"], m["x"]) for m in members]

    def batch():
        rc, st = N.pcg_batch(jobs, **kw)
        if rc:
            raise RuntimeError(f"ensemble solve failed: {rc}")
        for j in jobs:
            j[0].sync()
        return st
    ms_batch, stb = timed(batch, args.steps, args.warmup)
    log(f"64 streams (mlp_pcg_batch), FP64: {ms_batch:.2f} ms ({len(jobs) / ms_batch * 1e3:.0f} solves/s)")
    # each system's stacked FP64 solution against its own solve on its own stream (bitwise by construction)
    same = True
    err32 = 0.0
    for g, xs64, xs32 in zip(stacks, x64, x32):
        for k, i in enumerate(g["idx"]):
            xi = members[i]["x"].get()
            o0, o1 = int(g["off"][k]), int(g["off"][k + 1])
            same = same and np.array_equal(xs64[o0:o1].view(np.uint64), xi.view(np.uint64))
            err32 = max(err32, float(np.linalg.norm(xs32[o0:o1] - xi) / np.linalg.norm(xi)))
    t0 = time.perf_counter()
    for ctx, A, B, b, x in jobs:
        N.pcg(ctx, A, B, b, x, **kw)
    for j in jobs:
        j[0].sync()
    ms_seq = 1e3 * (time.perf_counter() - t0)
    log(f"one after another: {ms_seq:.2f} ms ({len(jobs) / ms_seq * 1e3:.0f} solves/s); stacked FP64 solutions "
        f"{'equal' if same else 'DIFFER FROM'} the single solves bit for bit; FP32 within {err32:.1e} of FP64")
    out = {
        "metric": "ensemble PCG solves per second (64 SPE10-shaped 60x220xnz systems, 1e-8 rel-res, FP64 outer, "
                  "FP32 preconditioner)",
        "value": round(len(members) / ms32 * 1e3, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms32, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {"workload": f"{len(members)} SPE10-shaped 60x220xnz systems (nz from 5/10/20/40/85, {cells} cells "
                               "in total), reference spe.cc assembly c=1, uh_cheb sigma=3 ncheb=2, PCG x0=0 to "
                               f"{args.eps:g}; systems grouped by hierarchy depth, each group one block-diagonal "
                               "system solved by mlp_pcg_ensemble (every launch covers the group), groups on "
                               "concurrent streams",
                   "nz": [nz for nz, _ in spec],
                   "groups": [{"depth": g["depth"], "systems": len(g["idx"]), "rows": int(g["off"][-1]),
                               "tuned_fp32": g["B32"].tune, "tuned_fp64": g["B64"].tune} for g in stacks]},
        "ms_stacked_fp64": round(ms64, 3),
        "ms_streams_fp64": round(ms_batch, 3),
        "ms_one_after_another": round(ms_seq, 3),
        "speedup_stacked_vs_streams_fp64": round(ms_batch / ms64, 3),
        "speedup_concurrent": round(ms_seq / ms32, 3),
        "stacked_fp64_bitwise_equal_single": bool(same),
        "fp32_rel_err_vs_fp64": err32,
        "iterations": it32,
        "worst_true_rel_res": worst32,
        "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(out), flush=True)
    for g in stacks:
        g["B32"].close()
        g["B64"].close()
        g["A"].close()
    for m in members:
        m["B"].close()
        m["A"].close()
    del jobs
    for g in stacks:
        del g["b"], g["x"]
        g["ctx"].close()
    for m in members:
        del m["b"], m["x"]
        m["ctx"].close()
    return out


def _rhs_cpu_job(item):
    """Right-hand sides of the 60x220xnz system on one host core: oracle construction, then one timed PCG solve
    per right-hand side."""
    nz, seeds, eps = item
    import oracle as O
    from ml_prec_amd import spe_synth as S
    s = S.spe10_system(kind="spe", c=1.0, nz=nz)
    P = O.cheb_prec(s.ia, s.ja, s.a)
    out = []
    for k in seeds:
        f = S.rhs(s.n, S.BASE_SEED + 500 + k)
        t0 = time.perf_counter()
        O.pcg(s.ia, s.ja, s.a, f, P, eps=eps, x0=np.zeros(s.n), maxit=10000)
        out.append(time.perf_counter() - t0)
    return out


def rhs_cpu_baseline(nz, R, eps, procs):
    """The C oracle on the R right-hand sides, spread over `procs` processes (each builds the hierarchy once;
    spawned before this process uses the GPU); aggregate = R over the summed solve time divided by procs."""
    import multiprocessing as mp
    procs = min(procs, R)
    parts = [list(range(k, R, procs)) for k in range(procs)]
    t0 = time.perf_counter()
    with mp.get_context("spawn").Pool(procs) as pool:
        res = pool.map(_rhs_cpu_job, [(nz, part, eps) for part in parts], chunksize=1)
    wall = time.perf_counter() - t0
    tsum = sum(sum(r) for r in res)
    return {"value": round(R / (tsum / procs), 2), "unit": "solves/s", "cores": procs, "kind": "port",
            "host": host_cpu(),
            "sample": f"all {R} right-hand sides, one oracle PCG solve each in {procs} processes (construction "
                      f"excluded); summed single-core solve time {tsum:.2f} s, wall time with construction {wall:.1f} s"}


def ensemble_rhs(N, S, args):
    """Many right-hand sides: one 60x220x85 system with COUNT seeded right-hand sides.  Batched: one mlp_pcg_multi call
    whose every launch covers all right-hand sides (the reported value).  For comparison: the same solves as
    concurrent jobs on clones of the preconditioner (mlp_pcg_batch) and one after another (mlp_pcg)."""
    R = args.ensemble_rhs
    cpu = None
    if not args.no_cpu:
        cpu = rhs_cpu_baseline(args.nz, R, args.eps, max(1, min(16, os.cpu_count() or 1)))
        log(f"CPU oracle: {cpu['value']:.1f} solves/s aggregate ({cpu['sample']})")
    s = S.spe10_system(kind="spe", c=1.0, nz=args.nz)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    ctx0 = N.Context(0)
    A = N.DevMatrix(ctx0, s.ia, s.ja, s.a)
    B0 = N.DevPrec(ctx0, A, H)
    fs = np.stack([S.rhs(s.n, S.BASE_SEED + 500 + k) for k in range(R)])
    kw = dict(eps=args.eps, maxit=10000, x0_mode=N.X0_ZERO, check_every=8, use_graph=bool(args.graph))

    def timed(run):
        for _ in range(max(1, args.warmup)):
            run()
        ctx0.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = run()
        ctx0.sync()
        return 1e3 * (time.perf_counter() - t0) / args.steps, out

    # batched right-hand sides
    bm = ctx0.to_device(fs.ravel())
    xm = ctx0.array(R * s.n)

    def multi():
        rc, st = N.pcg_multi(ctx0, A, B0, R, bm, xm, **kw)
        if rc:
            raise RuntimeError(f"batched solve failed: {rc}")
        return st

    ms_multi, st_multi = timed(multi)
    # the same batch with the FP32 hierarchy in FP64 flexible PCG (true residual every 20 iterations)
    B32 = N.DevPrec(ctx0, A, H, precision=N.MLP_FP32)
    xm32 = ctx0.array(R * s.n)

    def multi32():
        rc, st = N.pcg_multi(ctx0, A, B32, R, bm, xm32, flexible=True, true_res_every=20, **kw)
        if rc:
            raise RuntimeError(f"batched mixed solve failed: {rc}")
        return st

    ms_multi32, st_multi32 = timed(multi32)
    B32.close()
    # concurrent jobs on clones
    jobs = []
    for k in range(R):
        c = N.Context(0)
        jobs.append((c, A, B0.clone(c), c.to_device(fs[k]), c.array(s.n)))

    def batch():
        rc, st = N.pcg_batch(jobs, **kw)
        if rc:
            raise RuntimeError(f"solve failed: {rc}")
        for j in jobs:
            j[0].sync()
        return st

    ms_batch, st_batch = timed(batch)
    t0 = time.perf_counter()
    for c, Aj, Bj, b, x in jobs:
        N.pcg(c, Aj, Bj, b, x, **kw)
    for j in jobs:
        j[0].sync()
    ms_seq = 1e3 * (time.perf_counter() - t0)
    same = all(np.array_equal(xm.get()[k * s.n:(k + 1) * s.n], jobs[k][4].get()) for k in range(R))
    iters = [d["niter"] for d in st_multi]
    log(f"{R} right-hand sides, mixed (FP32 hierarchy): batched {ms_multi32:.2f} ms "
        f"({R / ms_multi32 * 1e3:.0f} solves/s), {ms_multi / ms_multi32:.2f}x the FP64 batch, worst true rel res "
        f"{max(d['true_rel_res'] for d in st_multi32):.2e}")
    log(f"{R} right-hand sides: batched {ms_multi:.2f} ms ({R / ms_multi * 1e3:.0f} solves/s), concurrent jobs "
        f"{ms_batch:.2f} ms ({R / ms_batch * 1e3:.0f} solves/s), one after another {ms_seq:.2f} ms, "
        f"iterations {min(iters)}..{max(iters)}, batched == single bitwise: {same}")
    print(json.dumps({
        "metric": "PCG solves per second, one 1.12M-cell system with many right-hand sides",
        "value": round(R / ms_multi * 1e3, 2), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_multi, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"SPE10-shaped 60x220x{args.nz}, {R} seeded right-hand sides, one uh_cheb hierarchy, "
                               "PCG x0=0 to 1e-8, all right-hand sides in every launch (mlp_pcg_multi)"},
        "ms_concurrent_jobs": round(ms_batch, 3), "ms_one_after_another": round(ms_seq, 3),
        "mixed_fp32_hierarchy": {"ms_per_step": round(ms_multi32, 3),
                                 "solves_per_s": round(R / ms_multi32 * 1e3, 2),
                                 "speedup_vs_fp64_batch": round(ms_multi / ms_multi32, 3),
                                 "worst_true_rel_res": max(d["true_rel_res"] for d in st_multi32)},
        "speedup_vs_one_after_another": round(ms_seq / ms_multi, 3), "bitwise_equal_to_single": same,
        "iterations": iters, "cpu_baseline": cpu}), flush=True)
    for c, Aj, Bj, b, x in jobs:
        Bj.close()
    B0.close()
    A.close()
    for j in jobs:
        j[0].close()
    ctx0.close()


def adjoint_cpu_baseline(S, shape, theta, b, obs, d, eps=1e-10):
    """One misfit gradient on one host core with the C oracle's construction and PCG solves (forward and
    adjoint to eps), the numpy assembly and the face-by-face gradient formula in numpy."""
    import oracle as O
    from ml_prec_amd.adjoint import KZ_RATIO, misfit
    t0 = time.perf_counter()
    kx = np.exp(theta)
    ia, ja, a = S.assemble(kx.reshape(shape), kind="spe", c=1.0)
    P = O.cheb_prec(ia, ja, a)
    n = kx.size
    p, st_f = O.pcg(ia, ja, a, b, P, eps=eps, x0=np.zeros(n), maxit=20000)
    _, g = misfit(p, obs, d)
    lam, st_a = O.pcg(ia, ja, a, g, P, eps=eps, x0=np.zeros(n), maxit=20000)
    k3, p3, l3 = kx.reshape(shape), p.reshape(shape), lam.reshape(shape)
    grad = np.zeros(shape)
    for axis, fac, kr in ((2, 1 / S.HX**2, 1.0), (1, 1 / S.HY**2, 1.0), (0, 1 / S.HZ**2, KZ_RATIO)):
        lo, hi = [slice(None)] * 3, [slice(None)] * 3
        lo[axis], hi[axis] = slice(None, -1), slice(1, None)
        ka, kb = kr * k3[tuple(lo)], kr * k3[tuple(hi)]
        T = 2.0 / (1.0 / ka + 1.0 / kb) * fac
        w = (l3[tuple(lo)] - l3[tuple(hi)]) * (p3[tuple(lo)] - p3[tuple(hi)])
        grad[tuple(lo)] -= T * T / (2 * fac * ka) * w
        grad[tuple(hi)] -= T * T / (2 * fac * kb) * w
    t = time.perf_counter() - t0
    return {"value": round(1e3 * t, 1), "unit": "ms per gradient", "cores": 1, "kind": "port", "host": host_cpu(),
            "sample": f"one gradient: numpy assembly, oracle construction, forward and adjoint oracle PCG to {eps:g} "
                      f"({st_f['niter']} and {st_a['niter']} iterations), numpy gradient formula"}, grad.ravel()


def adjoint(N, S, args, emit=True):
    """Adjoint gradient: L = sum over 200 observed cells of (p_i - d_i)^2 with d = p(true field) + N(0, 0.01|p_i|);
    theta = ln kx of a second field; one gradient = forward solve + adjoint solve (1e-10) + gradient kernel.
    Checked against central differences (h = 1e-6) along --fd-dirs seeded directions."""
    from ml_prec_amd.adjoint import SpeAdjointModel, misfit
    nz = args.nz
    shape = (nz, 220, 60)
    n = 60 * 220 * nz
    g = np.random.default_rng(S.BASE_SEED + 4)
    theta_true = np.log(S.spe10_field(60, 220, nz, S.BASE_SEED + 1).ravel())
    theta = np.log(S.spe10_field(60, 220, nz, S.BASE_SEED + 40).ravel())
    b = S.rhs(n, S.BASE_SEED + 1)
    ctx = N.Context(0)
    model = SpeAdjointModel(ctx, shape, b, eps=1e-10)
    obs = g.choice(n, 200, replace=False)
    p_true = model.forward(theta_true)
    d = p_true[obs] + g.normal(0.0, 1.0, obs.size) * 0.01 * np.abs(p_true[obs])
    times, phases = [], []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        p = model.forward(theta)
        L, gp = misfit(p, obs, d)
        grad = model.vjp(gp)
        times.append(time.perf_counter() - t0)
        phases.append(dict(model.phases))
    fw, adj = model.stats[-2][1], model.stats[-1][1]
    ph = {k: round(float(np.median([q[k] for q in phases])), 2) for k in phases[0]}
    log(f"adjoint gradient: {1e3 * np.median(times):.1f} ms (matrix assembled on the device {ph['matrix']:.1f} ms, "
        f"hierarchy built on the device {ph['construction']:.1f} ms, host planning + upload {ph['plan_upload']:.1f} "
        f"ms; forward {fw['niter']} and adjoint {adj['niter']} PCG iterations, solve times {fw['t_solve_ms']:.2f} + "
        f"{adj['t_solve_ms']:.2f} ms)")
    errs = []
    h = 1e-6
    for k in range(args.fd_dirs):
        v = np.random.default_rng(S.BASE_SEED + 100 + k).standard_normal(n)
        Lp, _ = misfit(model.forward(theta + h * v), obs, d)
        Lm, _ = misfit(model.forward(theta - h * v), obs, d)
        fd = (Lp - Lm) / (2 * h)
        errs.append(abs(fd - grad @ v) / abs(fd))
    if errs:
        log(f"central differences along {args.fd_dirs} directions: max rel. difference {max(errs):.2e}, "
            f"median {np.median(errs):.2e}")
    cpu = None
    if not args.no_cpu:
        cpu, gcpu = adjoint_cpu_baseline(S, shape, theta, b, obs, d)
        cpu["rel_diff_vs_gpu"] = float(np.linalg.norm(gcpu - grad) / np.linalg.norm(grad))
        log(f"CPU oracle: {cpu['value']:.0f} ms per gradient ({cpu['sample']}); gradients differ by "
            f"{cpu['rel_diff_vs_gpu']:.1e} relative")
    out = {
        "metric": "adjoint misfit gradient time at 1.12M cells (forward + adjoint PCG to 1e-10)",
        "value": round(1e3 * float(np.median(times)), 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": 0,
        "ms_per_step": round(1e3 * float(np.median(times)), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"SPE10-shaped 60x220x{nz} ({n} cells), theta = ln kx, 200 observations, "
                               "uh_cheb PCG forward and adjoint to 1e-10"},
        "device_solve_ms": {"forward": fw["t_solve_ms"], "adjoint": adj["t_solve_ms"]},
        "phases_ms": dict(ph, **{"what": "host wall time per forward: device assembly of the matrix, device "
                                         "construction of the hierarchy (level arrays copied to the host), host "
                                         "planning of the apply programs and their upload"}),
        "iterations": {"forward": fw["niter"], "adjoint": adj["niter"]},
        "misfit": L,
        "fd_check": {"directions": args.fd_dirs, "h": h, "max_rel_diff": max(errs) if errs else None,
                     "median_rel_diff": float(np.median(errs)) if errs else None},
        "cpu_baseline": cpu,
    }
    if emit:
        print(json.dumps(out), flush=True)
    model.close()
    ctx.close()
    return out


MSPLIT_CFG = dict(q=(0.8,), niters=(5,), order="simple_1", degree_max=2)   # project/msplit.cfg:7-77
GMRES_CAP = 145       # five restart cycles of 29 steps
SOLVER_CASES = (
    # name, matrix, outer solver, preconditioner, iteration cap (None: solve to eps)
    ("pcg_uh_cheb", "sym", "pcg", "uh_cheb", None),
    ("pcg_multi_split", "sym", "pcg", "multi_split", None),
    # FP32 hierarchy inside FP64 flexible PCG, true residual every 20 iterations
    ("pcg_mixed_uh_cheb", "sym", "pcg_mixed", "uh_cheb", None),
    ("pcg_mixed_multi_split", "sym", "pcg_mixed", "multi_split", None),
    # FP64 hierarchy in the same flexible PCG (multi_split is not a symmetric preconditioner)
    ("pcg_flex_multi_split", "sym", "pcg_flex", "multi_split", None),
    ("gmres_uh_cheb", "sym", "gmres", "uh_cheb", None),
    ("cheb_solver_uh_cheb", "sym", "cheb", "uh_cheb", None),
    ("simple_solver_multi_split", "sym", "simple", "multi_split", None),
    # GMRES(30) with GS alone or GS + multi_split stalls on the 1e6-contrast nonsymmetric system (relative
    # residuals around 1e-5 and 5e-4 after five cycles, in the oracle too): timed over five cycles
    ("gmres_gs", "nonsym", "gmres", "gs", GMRES_CAP),
    ("gmres_composite", "nonsym", "gmres", "composite", GMRES_CAP),
    ("gmres_sym", "nonsym", "gmres", "sym", None),
)


def cpu_case_baseline(O, s, a, case, gpu_niter, budget_s):
    """The C oracle (one thread) on the same case: a full solve to 1e-8 when the estimate (GPU iteration count x
    one timed preconditioner apply, x2.5 for GMRES's orthogonalisation, x1.3 otherwise) fits `budget_s`;
    otherwise only the preconditioner applies are priced (GPU iteration count x one apply), a lower bound of the
    CPU solve, marked as such."""
    _, _, solver, prec, cap = case
    t0 = time.perf_counter()
    if prec == "uh_cheb":
        P = O.cheb_prec(s.ia, s.ja, a)
    elif prec == "multi_split":
        P = O.msplit_prec(s.ia, s.ja, a, **MSPLIT_CFG)
    elif prec == "gs":
        P = O.gs_prec(s.ia, s.ja, a)
    elif prec == "composite":
        P = O.composite_prec(O.gs_prec(s.ia, s.ja, a), O.msplit_prec(s.ia, s.ja, a, **MSPLIT_CFG))
    else:
        P = O.sym_prec(s.ia, s.ja, a)
    t_const = time.perf_counter() - t0
    t0 = time.perf_counter()
    P.apply(s.b)
    t_apply = time.perf_counter() - t0
    est = t_apply * gpu_niter * (2.5 if solver == "gmres" else 1.3)
    rec = {"apply_ms": round(1e3 * t_apply, 3), "construction_s": round(t_const, 3), "cores": 1, "kind": "port",
           "host": host_cpu()}
    if est > budget_s:
        rec.update(ms_per_solve=round(1e3 * t_apply * gpu_niter, 2), niter=None,
                   sample=f"one preconditioner apply x {gpu_niter} iterations (lower bound; a full solve is "
                          f"estimated at {est:.0f} s)")
        return rec, None
    kw = dict(eps=1e-8, x0=np.zeros(s.n), maxit=cap or 20000, check=cap is None)
    t0 = time.perf_counter()
    if solver in ("pcg", "pcg_mixed", "pcg_flex"):        # the oracle's PCG (no FP32 hierarchy, no flexible beta)
        x, st = O.pcg(s.ia, s.ja, a, s.b, P, **kw)
    elif solver == "gmres":
        x, st = O.gmres(s.ia, s.ja, a, s.b, P, **kw)
    elif solver == "cheb":
        li = P.level_info(0)
        x, st = O.cheb_solver(s.ia, s.ja, a, s.b, P, li["lmin"], li["lmax"], **kw)
    else:
        x, st = O.simple_solver(s.ia, s.ja, a, s.b, P, **kw)
    rec.update(ms_per_solve=round(1e3 * (time.perf_counter() - t0), 2), niter=int(st["niter"]),
               sample="one full solve" if cap is None else f"one solve capped at {cap} iterations")
    return rec, x


def solvers(N, S, args):
    """Every outer solver with the preconditioners it is used with, on the 60x220x85 system (and, for GMRES
    with GS / Composite / Sym, on the reference's nonsymmetric variant of it, spe.cc:138-182 with tau = 0.3):
    device solve time to 1e-8 from x0 = 0, iterations, true residual, preconditioner apply time, and the
    single-core C oracle on the same case.  Prints one JSON line with a record per case."""
    s = S.spe10_system(kind="spe", c=1.0, nz=args.nz)
    a_ns = S.make_nonsymmetric(s.ia, s.ja, s.a, 0.3)
    cases = [c for c in SOLVER_CASES if not args.case or c[0] in args.case.split(",")]
    if not cases:
        raise SystemExit(f"--case: none of {args.case!r} in {[c[0] for c in SOLVER_CASES]}")
    if not args.no_cpu:
        import oracle as O
    ctx = N.Context(0)
    out = []
    for case in cases:
        name, mat, solver, prec, cap = case
        a = s.a if mat == "sym" else a_ns
        t0 = time.perf_counter()
        A = N.DevMatrix(ctx, s.ia, s.ja, a)
        H = None
        precision = N.MLP_FP32 if solver == "pcg_mixed" else N.MLP_FP64
        if prec == "uh_cheb":
            H = N.build_uh_cheb(s.ia, s.ja, a)
            B = N.DevPrec(ctx, A, H, precision=precision)
        elif prec == "multi_split":
            H = N.build_multi_split(s.ia, s.ja, a, **MSPLIT_CFG)
            B = N.DevPrec(ctx, A, H, precision=precision)
        elif prec == "gs":
            B = N.DevPrec(ctx, A, kind=N.GS)
        elif prec == "composite":
            H = N.build_multi_split(s.ia, s.ja, a, **MSPLIT_CFG)
            B = N.DevPrec(ctx, A, H, kind=N.COMPOSITE)
        else:
            B = N.DevPrec(ctx, A, kind=N.SYM)
        ctx.sync()
        t_setup = time.perf_counter() - t0
        b = ctx.to_device(s.b)
        x = ctx.array(s.n)
        r = ctx.array(s.n)
        kw = dict(eps=args.eps, maxit=cap or 20000, x0_mode=N.X0_ZERO, raise_on_error=cap is None)
        if solver == "pcg":
            run = lambda: N.pcg(ctx, A, B, b, x, check_every=8, use_graph=bool(args.graph), **kw)
        elif solver in ("pcg_mixed", "pcg_flex"):
            run = lambda: N.pcg(ctx, A, B, b, x, check_every=8, use_graph=bool(args.graph), flexible=True,
                                true_res_every=20, **kw)
        elif solver == "gmres":
            run = lambda: N.gmres(ctx, A, B, b, x, **kw)
        elif solver == "cheb":
            li = H.level_info(0)
            run = lambda: N.cheb_solver(ctx, A, B, li["lmin"], li["lmax"], b, x, check_every=8,
                                        use_graph=bool(args.graph), **kw)
        else:
            run = lambda: N.simple_solver(ctx, A, B, b, x, check_every=8, use_graph=bool(args.graph), **kw)
        for _ in range(max(1, args.warmup)):
            run()
        ctx.sync()
        times = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            st = run()
            ctx.sync()
            times.append(1e3 * (time.perf_counter() - t0))
            if st["rc"] not in (0, 5):                   # 5: the iteration cap of a capped case
                raise RuntimeError(f"{name}: solve failed with status {st['rc']}")
        A.residual(b, x, r)
        true_res = N.dnrm2(ctx, r) / N.dnrm2(ctx, b)
        f = ctx.to_device(np.random.default_rng(5).standard_normal(s.n))
        apply_ms = B.time_apply(f, r, reps=10)
        rec = {"case": name, "matrix": "SPE10-shaped 60x220x%d%s" % (args.nz, "" if mat == "sym" else
                                                                      ", nonsymmetric (tau 0.3)"),
               "solver": solver, "preconditioner": prec, "ms_per_solve": round(float(np.median(times)), 3),
               "niter": int(st["niter"]), "iteration_cap": cap, "true_rel_res": true_res,
               "apply_us": round(1e3 * apply_ms, 1), "setup_s": round(t_setup, 3)}
        log(f"{name}: {rec['ms_per_solve']:.2f} ms/solve, {rec['niter']} iterations, true rel res {true_res:.2e}, "
            f"apply {rec['apply_us']:.1f} us, setup {t_setup:.2f} s")
        if not args.no_cpu:
            cpu, xo = cpu_case_baseline(O, s, a, case, rec["niter"], args.cpu_budget)
            rec["cpu_baseline"] = cpu
            rec["speedup_vs_cpu"] = round(cpu["ms_per_solve"] / rec["ms_per_solve"], 1)
            if xo is not None:
                rec["rel_diff_vs_cpu"] = float(np.linalg.norm(x.get() - xo) / np.linalg.norm(xo))
            log(f"  oracle: {cpu['ms_per_solve']:.1f} ms/solve ({cpu['sample']}), iterations {cpu['niter']}"
                f"{', solution rel diff %.1e' % rec['rel_diff_vs_cpu'] if xo is not None else ''}")
        out.append(rec)
        B.close()
        A.close()
        del H, b, x, r, f
    print(json.dumps({
        "metric": "solve time to 1e-8 rel-res @1.12M cells per outer solver and preconditioner",
        "value": round(sum(r["ms_per_solve"] for r in out), 3), "unit": "ms (sum over cases)", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"SPE10-shaped 60x220x{args.nz}, reference spe.cc assembly c=1, x0=0, eps "
                               f"{args.eps:g}; multi_split with project/msplit.cfg settings"},
        "cases": out}), flush=True)
    ctx.close()


def level_descs(H, first, s):
    """Levels first..last of host hierarchy H as mlp_level_desc dicts (a standalone hierarchy whose top level is
    `first`) and the top level's matrix (ia, ja, a)."""
    out = []
    prev = None
    for k in range(H.nlevels):
        info, arr = H.level_arrays(k)
        if k >= first:
            d = dict(N=info["N"], M=info["M"], Md=info["Md"], nit=info["nit"], lmin=info["lmin"], lmax=info["lmax"],
                     eps=info["eps"], map=arr["map"], L_ia=arr["L_ia"], L_ja=arr["L_ja"], L_a=arr["L_a"],
                     U_ia=arr["U_ia"], U_ja=arr["U_ja"], U_a=arr["U_a"], dval=arr["dval"])
            if k > first:
                d.update(A_ia=prev["A_ia"], A_ja=prev["A_ja"], A_a=prev["A_a"])
            out.append(d)
        if k + 1 == first:
            top = (arr["A_ia"], arr["A_ja"], arr["A_a"])
        prev = arr
    if first == 0:
        top = (s.ia, s.ja, s.a)
    return out, top


def error_budget(N, S, args):
    """Error of the FP32 hierarchy level by level: for each level l the hierarchy from l down to the coarsest
    (whose solve is FP64 in both) is uploaded in FP64 and in FP32 and applied to the same seeded vectors; the
    relative difference of the two results is the error the FP32 storage and arithmetic of levels l.. add.  The
    same for the BF16 variant (level l's Chebyshev vectors in BF16 on top of FP32).  Also the apply times and, for
    the whole hierarchy, the effect on PCG to 1e-8 and 1e-10 (iterations, solution error against FP64)."""
    s = S.spe10_system(kind="spe", c=1.0, nz=args.nz)
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    ctx = N.Context(0)
    rows = []
    g = np.random.default_rng(11)
    for l in range(H.nlevels):
        levels, (ia, ja, a) = level_descs(H, l, s)
        A = N.DevMatrix(ctx, ia, ja, a)
        B64 = N.DevPrec.upload(ctx, A, N.UH_CHEB, levels, N.MLP_FP64)
        B32 = N.DevPrec.upload(ctx, A, N.UH_CHEB, levels, N.MLP_FP32)
        n = len(ia) - 1
        errs = []
        y64, y32 = ctx.array(n), ctx.array(n)
        for k in range(4):
            f = ctx.to_device(g.standard_normal(n))
            B64.apply(f, y64)
            B32.apply(f, y32)
            a64, a32 = y64.get(), y32.get()
            errs.append(float(np.linalg.norm(a32 - a64) / np.linalg.norm(a64)))
        f = ctx.to_device(g.standard_normal(n))
        t64 = min(B64.time_apply(f, y64, reps=20) for _ in range(3))
        t32 = min(B32.time_apply(f, y32, reps=20) for _ in range(3))
        row = {"level": l, "N": n, "fp32_rel_err_max": max(errs), "fp32_rel_err_mean": float(np.mean(errs)),
               "apply_us_fp64": round(t64 * 1e3, 2), "apply_us_fp32": round(t32 * 1e3, 2),
               "launches": B64.launches_per_apply()}
        msg = ""
        if l + 1 < H.nlevels:
            # BF16 on this level's Chebyshev vectors (the variant applied to the hierarchy from level l)
            Bbf = N.DevPrec.upload(ctx, A, N.UH_CHEB, levels, N.MLP_FP32_BF16)
            ebf = []
            ybf = ctx.array(n)
            for k in range(4):
                f = ctx.to_device(g.standard_normal(n))
                B64.apply(f, y64)
                Bbf.apply(f, ybf)
                a64 = y64.get()
                ebf.append(float(np.linalg.norm(ybf.get() - a64) / np.linalg.norm(a64)))
            tbf = min(Bbf.time_apply(f, ybf, reps=20) for _ in range(3))
            row.update(bf16_rel_err_max=max(ebf), bf16_rel_err_mean=float(np.mean(ebf)),
                       apply_us_fp32_bf16=round(tbf * 1e3, 2))
            msg = f"; BF16 Chebyshev vectors rel err {max(ebf):.3e}, apply {tbf * 1e3:.1f} us"
            Bbf.close()
        log(f"levels {l}..{H.nlevels - 1}: N={n} FP32 operator rel err {max(errs):.3e}; apply FP64 "
            f"{t64 * 1e3:.1f} us, FP32 {t32 * 1e3:.1f} us" + msg)
        rows.append(row)
        B32.close()
        B64.close()
        A.close()
    # the whole hierarchy inside PCG, to 1e-8 and 1e-10: iterations and error against the FP64 solution
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B64 = N.DevPrec(ctx, A, H, precision=N.MLP_FP64)
    B32 = N.DevPrec(ctx, A, H, precision=N.MLP_FP32)
    Bbf = N.DevPrec(ctx, A, H, precision=N.MLP_FP32_BF16)
    b = ctx.to_device(s.b)
    x = ctx.array(s.n)
    pcg = {}
    for eps in (1e-8, 1e-10):
        e = {}
        st64 = N.pcg(ctx, A, B64, b, x, eps=eps, maxit=10000, use_graph=True, check_every=8)
        x64 = x.get()
        e["niter_fp64"] = int(st64["niter"])
        for name, B in (("mixed", B32), ("mixed_bf16", Bbf)):
            st = N.pcg(ctx, A, B, b, x, eps=eps, maxit=10000, flexible=True, true_res_every=20, use_graph=True,
                       check_every=8)
            xm = x.get()
            e[f"niter_{name}"] = int(st["niter"])
            e[f"true_rel_res_{name}"] = st["true_rel_res"]
            e[f"{name}_rel_err_vs_fp64"] = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
        pcg[f"{eps:g}"] = e
        log(f"PCG to {eps:g}: FP64 {e['niter_fp64']} it, mixed {e['niter_mixed']} it (vs FP64 "
            f"{e['mixed_rel_err_vs_fp64']:.3e}), mixed+BF16 {e['niter_mixed_bf16']} it (vs FP64 "
            f"{e['mixed_bf16_rel_err_vs_fp64']:.3e})")
    print(json.dumps({"error_budget": rows, "pcg": pcg,
                      "workload": f"SPE10-shaped {s.nx}x{s.ny}x{s.nz}, uh_cheb sigma 3 ncheb 2, "
                                  f"{H.nlevels} levels"}), flush=True)
    for o in (Bbf, B32, B64, A):
        o.close()
    del b, x
    ctx.close()


def apply_traffic_probe(N, s, reps):
    """Preconditioner applies (FP64 and FP32 hierarchies, `reps` each, launched from the host) and the same
    1 GiB calibration kernel as the SpMV probe, for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs: the level
    kernels' counters summed per apply give the DRAM bytes one apply moves."""
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    ctx = N.Context(0)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    f = ctx.to_device(np.random.default_rng(2).standard_normal(s.n))
    z = ctx.array(s.n)
    out = {"apply_traffic_probe": True, "reps": reps}
    for prec, name in ((N.MLP_FP64, "fp64"), (N.MLP_FP32, "fp32")):
        B = N.DevPrec(ctx, A, H, precision=prec)
        ms = B.time_apply(f, z, reps=reps)
        out[name] = {"algorithmic_bytes_per_apply": B.bytes_per_apply(), "us_per_apply": round(ms * 1e3, 2)}
        log(f"{name}: {reps} applies, {ms * 1e3:.1f} us each")
        ctx.sync()
        B.close()
    big = ctx.array(1 << 27)
    for _ in range(3):
        N.dscal(ctx, 0.5, big)
    ctx.sync()
    log("calibration: 3 launches of x = 0.5 x over 2^27 doubles (1 GiB read + 1 GiB written each)")
    print(json.dumps(out), flush=True)
    A.close()
    del f, z, big
    ctx.close()


def spmv_traffic_probe(N, s, reps):
    """Cold SpMV launches plus a calibration kernel of known traffic: x = 0.5 x
    over 2^27 doubles reads and writes exactly 1 GiB each.  Used under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to measure HBM bytes per launch."""
    ctx = N.Context(0)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    xv = ctx.to_device(np.random.default_rng(0).standard_normal(s.n))
    yv = ctx.array(s.n)
    ms = A.time_spmv(xv, yv, reps=reps, flush_bytes=1 << 30)
    log(f"SpMV cold {ms * 1e3:.2f} us over {reps} launches")
    big = ctx.array(1 << 27)
    for _ in range(3):
        N.dscal(ctx, 0.5, big)
    ctx.sync()
    log("calibration: 3 launches of x = 0.5 x over 2^27 doubles (1 GiB read + 1 GiB written each)")
    _, _, dev_bytes = A.info()
    print(json.dumps({"spmv_traffic_probe": True, "n": s.n, "nnz": s.nnz, "device_bytes": dev_bytes,
                      "spmv_us_cold": round(ms * 1e3, 3)}), flush=True)
    A.close()
    del xv, yv, big
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the ensemble and adjoint summaries of the default run")
    ap.add_argument("--spmv-reps", type=int, default=200)
    ap.add_argument("--spmv-warm-reps", type=int, default=None, help="0 skips the cache-warm SpMV timing")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--nz", type=int, default=85)
    ap.add_argument("--ensemble", type=int, default=0, metavar="COUNT",
                    help="an ensemble instead: COUNT independent systems (nz drawn from 5/10/20/40/85) solved "
                         "concurrently; prints its own JSON line")
    ap.add_argument("--ensemble-rhs", type=int, default=0, metavar="COUNT",
                    help="the 60x220x85 system with COUNT right-hand sides instead, one shared hierarchy")
    ap.add_argument("--adjoint", action="store_true",
                    help="the misfit gradient by the adjoint solve instead, checked by central differences")
    ap.add_argument("--fd-dirs", type=int, default=20)
    ap.add_argument("--solvers", action="store_true",
                    help="every outer solver / preconditioner pair at 1.12M cells instead; prints its own JSON line")
    ap.add_argument("--case", default="", help="with --solvers: comma-separated case names to run")
    ap.add_argument("--cpu-budget", type=float, default=25.0,
                    help="with --solvers: estimated seconds of single-core oracle solve allowed per case")
    ap.add_argument("--error-budget", action="store_true",
                    help="the FP32 hierarchy's error level by level against FP64 instead; prints its own JSON line")
    ap.add_argument("--apply-traffic", action="store_true",
                    help="only run preconditioner applies and a 1 GiB scale kernel (for HBM counter runs), then exit")
    ap.add_argument("--spmv-traffic", action="store_true",
                    help="only run cold SpMV launches and a 1 GiB scale kernel (for HBM counter runs), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (each system is solved on one "
              "device; there is no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    from ml_prec_amd import native as N
    from ml_prec_amd import spe_synth as S

    if args.ensemble:
        ensemble(N, S, args)
        return
    if args.adjoint:
        adjoint(N, S, args)
        return
    if args.ensemble_rhs:
        ensemble_rhs(N, S, args)
        return
    if args.solvers:
        solvers(N, S, args)
        return
    if args.error_budget:
        error_budget(N, S, args)
        return

    t0 = time.perf_counter()
    s = S.spe10_system(kind="spe", c=1.0, nz=args.nz)
    log(f"system {s.nx}x{s.ny}x{s.nz}: n={s.n} nnz={s.nnz} generated in {time.perf_counter() - t0:.2f}s")
    if args.spmv_traffic:
        spmv_traffic_probe(N, s, args.spmv_reps)
        return
    if args.apply_traffic:
        apply_traffic_probe(N, s, 20)
        return

    t0 = time.perf_counter()
    H = N.build_uh_cheb(s.ia, s.ja, s.a)
    t_const = time.perf_counter() - t0
    levels = [H.level_info(l) for l in range(H.nlevels)]
    log(f"hierarchy: {H.nlevels} levels, N = {[l['N'] for l in levels]}, built in {t_const:.2f}s")

    ctx = N.Context(0)
    A = N.DevMatrix(ctx, s.ia, s.ja, s.a)
    B64 = N.DevPrec(ctx, A, H, precision=N.MLP_FP64)
    B32 = N.DevPrec(ctx, A, H, precision=N.MLP_FP32)
    b = ctx.to_device(s.b)
    x = ctx.array(s.n)

    def solve_fp64():
        return N.pcg(ctx, A, B64, b, x, eps=args.eps, maxit=10000, use_graph=bool(args.graph), check_every=8)

    def solve_mixed():
        return N.pcg(ctx, A, B32, b, x, eps=args.eps, maxit=10000, flexible=True, true_res_every=20,
                     use_graph=bool(args.graph), check_every=8)

    for _ in range(args.warmup):
        solve_fp64()
        solve_mixed()
    ctx.sync()
    log("warmup done")

    ctx.sync()
    t0 = time.perf_counter()
    st64 = [solve_fp64() for _ in range(args.steps)]
    ctx.sync()
    ms64 = 1e3 * (time.perf_counter() - t0) / args.steps
    x64 = x.get()
    log(f"FP64: {ms64:.3f} ms/solve, {st64[-1]['niter']} iterations, true rel res {st64[-1]['true_rel_res']:.3e}")

    ctx.sync()
    t0 = time.perf_counter()
    stm = [solve_mixed() for _ in range(args.steps)]
    ctx.sync()
    msm = 1e3 * (time.perf_counter() - t0) / args.steps
    xm = x.get()
    err_mixed = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
    log(f"mixed: {msm:.3f} ms/solve, {stm[-1]['niter']} iterations, true rel res {stm[-1]['true_rel_res']:.3e}, "
        f"rel err vs FP64 GPU {err_mixed:.3e}")

    # the same solves to 1e-10, and the BF16 variant of the mixed solve (level 0's Chebyshev vectors in BF16)
    steps10 = max(3, args.steps // 2)

    def timed(fn):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        st = [fn() for _ in range(steps10)]
        ctx.sync()
        return 1e3 * (time.perf_counter() - t0) / steps10, st[-1]

    ms64_10, st64_10 = timed(lambda: N.pcg(ctx, A, B64, b, x, eps=1e-10, maxit=10000, use_graph=bool(args.graph),
                                           check_every=8))
    x64_10 = x.get()
    msm_10, stm_10 = timed(lambda: N.pcg(ctx, A, B32, b, x, eps=1e-10, maxit=10000, flexible=True,
                                         true_res_every=20, use_graph=bool(args.graph), check_every=8))
    err_m10 = float(np.linalg.norm(x.get() - x64_10) / np.linalg.norm(x64_10))
    log(f"to 1e-10: FP64 {ms64_10:.3f} ms ({st64_10['niter']} it), mixed {msm_10:.3f} ms ({stm_10['niter']} it), "
        f"mixed/FP64 {ms64_10 / msm_10:.3f}x, rel err vs FP64 {err_m10:.2e}")
    Bbf = N.DevPrec(ctx, A, H, precision=N.MLP_FP32_BF16)
    msbf, stbf = timed(lambda: N.pcg(ctx, A, Bbf, b, x, eps=args.eps, maxit=10000, flexible=True, true_res_every=20,
                                     use_graph=bool(args.graph), check_every=8))
    err_bf = float(np.linalg.norm(x.get() - x64) / np.linalg.norm(x64))
    applybf_ms = min(Bbf.time_apply(b, x, reps=20) for _ in range(3))
    log(f"mixed with BF16 Chebyshev vectors on level 0: {msbf:.3f} ms ({stbf['niter']} it), apply "
        f"{applybf_ms * 1e3:.1f} us, rel err vs FP64 {err_bf:.2e}")
    Bbf.close()

    # SpMV roofline: algorithmic bytes (w+4)*nnz + 4(N+1) + 2wN, cold L2 and Infinity Cache
    bytes_spmv = 12.0 * s.nnz + 4.0 * (s.n + 1) + 16.0 * s.n
    xv = ctx.to_device(np.random.default_rng(0).standard_normal(s.n))
    yv = ctx.array(s.n)
    spmv_cold_ms = A.time_spmv(xv, yv, reps=args.spmv_reps, flush_bytes=1 << 30)
    warm_reps = args.spmv_reps if args.spmv_warm_reps is None else args.spmv_warm_reps
    spmv_warm_ms = A.time_spmv(xv, yv, reps=warm_reps, flush_bytes=0) if warm_reps else None
    ach = bytes_spmv / (spmv_cold_ms * 1e-3) / 1e9
    log(f"SpMV: cold {spmv_cold_ms * 1e3:.2f} us ({ach:.0f} GB/s)" +
        (f", warm {spmv_warm_ms * 1e3:.2f} us ({bytes_spmv / (spmv_warm_ms * 1e-3) / 1e9:.0f} GB/s)"
         if spmv_warm_ms else ""))
    f = ctx.to_device(np.random.default_rng(1).standard_normal(s.n))
    z = ctx.array(s.n)
    # back-to-back applies launched from the host: the best of three runs of 20 after a warm-up run
    apply64_ms = min(B64.time_apply(f, z, reps=20) for _ in range(4))
    apply32_ms = min(B32.time_apply(f, z, reps=20) for _ in range(4))
    by64, by32 = B64.bytes_per_apply(), B32.bytes_per_apply()
    nl64, nl32 = B64.launches_per_apply(), B32.launches_per_apply()
    log(f"prec apply: FP64 {apply64_ms * 1e3:.1f} us ({by64 / apply64_ms / 1e6:.0f} GB/s alg), "
        f"FP32 {apply32_ms * 1e3:.1f} us ({by32 / apply32_ms / 1e6:.0f} GB/s alg)")

    cpu = None
    err_cpu = None
    if not args.no_cpu:
        log("timing the CPU oracle ...")
        cpu, xcpu = cpu_baseline(s, args.eps)
        err_cpu = float(np.linalg.norm(x64 - xcpu) / np.linalg.norm(xcpu))
        log(f"CPU oracle: {cpu['value']:.1f} ms/solve ({cpu['niter']} iterations); FP64 GPU vs CPU rel diff "
            f"{err_cpu:.3e}")

    out = {
        "metric": METRIC,
        "value": round(msm, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(msm, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": f"SPE10-shaped {s.nx}x{s.ny}x{s.nz} ({s.n} cells, {s.nnz} nnz), reference spe.cc "
                        "assembly c=1 on a seeded channelised field, uh_cheb sigma=3 ncheb=2, PCG x0=0 to "
                        f"{args.eps:g} relative residual",
            "levels": H.nlevels,
            "level_sizes": [l["N"] for l in levels],
            "mixed": "FP64 flexible PCG, FP32 hierarchy, true residual every 20 iterations",
        },
        "fp64_ms_per_solve": round(ms64, 4),
        "speedup_mixed_vs_fp64": round(ms64 / msm, 3),
        "niter_fp64": int(st64[-1]["niter"]),
        "niter_mixed": int(stm[-1]["niter"]),
        "true_rel_res_fp64": st64[-1]["true_rel_res"],
        "true_rel_res_mixed": stm[-1]["true_rel_res"],
        "mixed_rel_err_vs_fp64": err_mixed,
        "fp64_rel_diff_vs_cpu_oracle": err_cpu,
        "spmv_us_cold": round(spmv_cold_ms * 1e3, 3),
        "spmv_us_warm": round(spmv_warm_ms * 1e3, 3) if spmv_warm_ms else None,
        "spmv_gbs_warm_algorithmic": round(bytes_spmv / (spmv_warm_ms * 1e-3) / 1e9, 1) if spmv_warm_ms else None,
        "prec_apply_us": {"fp64": round(apply64_ms * 1e3, 2), "fp32": round(apply32_ms * 1e3, 2)},
        "prec_apply_gbs_algorithmic": {"fp64": round(by64 / apply64_ms / 1e6, 1),
                                       "fp32": round(by32 / apply32_ms / 1e6, 1)},
        "eps_1e-10": {"fp64_ms_per_solve": round(ms64_10, 4), "mixed_ms_per_solve": round(msm_10, 4),
                      "speedup_mixed_vs_fp64": round(ms64_10 / msm_10, 3), "niter_fp64": int(st64_10["niter"]),
                      "niter_mixed": int(stm_10["niter"]), "true_rel_res_mixed": stm_10["true_rel_res"],
                      "mixed_rel_err_vs_fp64": err_m10, "steps": steps10},
        "mixed_bf16": {"ms_per_solve": round(msbf, 4), "niter": int(stbf["niter"]),
                       "true_rel_res": stbf["true_rel_res"], "rel_err_vs_fp64": err_bf,
                       "prec_apply_us": round(applybf_ms * 1e3, 2),
                       "what": "FP32 hierarchy with level 0's Chebyshev vectors (F, u1, a1 u1, x2) in BF16, FP64 "
                               "flexible PCG"},
        "t_const_host_s": round(t_const, 3),
        # the dominant work of the headline (mixed) solve: the FP32-hierarchy preconditioner apply, one "launch"
        # being one apply (its level kernels back to back, HIP events on the context stream)
        "roofline": {
            "kernel": f"uh_cheb preconditioner apply, FP32 hierarchy: {nl32} launches per apply (the descent and "
                      f"ascent kernels of the level chain; {100 * apply32_ms * int(stm[-1]['niter']) / msm:.0f}% of "
                      "the mixed solve); achieved = algorithmic bytes per apply (sum over level visits of L/U sweeps, "
                      "level vectors, index arrays and the fused Chebyshev-residual SpMVs) / apply time; bound in "
                      "practice by dependent-load and sweep-depth latency, not by bandwidth",
            "bound": "hbm",
            "achieved": round(by32 / apply32_ms / 1e6, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(by32 / apply32_ms / 1e6 / HBM_PEAK_GBS, 4),
            "traffic": load_apply_traffic("fp32"),
            "algorithmic_bytes_per_launch": by32,
            "us_per_apply": round(apply32_ms * 1e3, 2),
        },
        "roofline_fp64_apply": {
            "kernel": f"the same apply with the FP64 hierarchy ({nl64} launches per apply)",
            "bound": "hbm",
            "achieved": round(by64 / apply64_ms / 1e6, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(by64 / apply64_ms / 1e6 / HBM_PEAK_GBS, 4),
            "traffic": load_apply_traffic("fp64"),
            "algorithmic_bytes_per_launch": by64,
            "us_per_apply": round(apply64_ms * 1e3, 2),
        },
        "roofline_spmv": {
            "kernel": "k_spmv_pairs, FP64, 1.12M rows in the symmetric seven-point layout (diagonal and three lower "
                      "values per row; upper values read from the neighbouring rows), two rows per thread with 16-byte "
                      "loads, launches rotating over 1 GiB of matrix copies so that the matrix streams from HBM",
            "bound": "hbm",
            "achieved_dram": round((load_traffic() or 0) / (spmv_cold_ms * 1e-3) / 1e9, 1) if load_traffic() else None,
            "frac_dram": round((load_traffic() or 0) / (spmv_cold_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
            if load_traffic() else None,
            "traffic": load_traffic(),
            "achieved_csr_equivalent": round(ach, 1),
            "frac_csr_equivalent": round(ach / HBM_PEAK_GBS, 4),
            "algorithmic_bytes_per_launch": bytes_spmv,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
        },
        "cpu_baseline": cpu,
    }
    for o in (B32, B64, A):
        o.close()
    del xv, yv, f, z, b, x
    ctx.close()
    if not args.no_extras:
        # the other workload configurations, summarised (their own modes print the full records)
        import copy
        a2 = copy.copy(args)
        a2.no_cpu, a2.steps, a2.warmup = True, 3, 1
        a2.ensemble = 64
        e = ensemble(N, S, a2, emit=False)
        out["ensemble_64"] = {"solves_per_s": e["value"], "ms_all_64": e["ms_per_step"],
                              "ms_stacked_fp64": e["ms_stacked_fp64"], "ms_streams_fp64": e["ms_streams_fp64"],
                              "ms_one_after_another": e["ms_one_after_another"],
                              "speedup_concurrent": e["speedup_concurrent"],
                              "stacked_fp64_bitwise_equal_single": e["stacked_fp64_bitwise_equal_single"],
                              "iterations": [min(e["iterations"]), max(e["iterations"])],
                              "worst_true_rel_res": e["worst_true_rel_res"], "groups": e["config"]["groups"],
                              "workload": e["config"]["workload"]}
        a2.steps, a2.fd_dirs = 3, 2
        d = adjoint(N, S, a2, emit=False)
        out["adjoint"] = {"ms_per_gradient": d["value"], "iterations": d["iterations"],
                          "phases_ms": {k: v for k, v in d["phases_ms"].items() if k != "what"},
                          "device_solve_ms": d["device_solve_ms"],
                          "fd_max_rel_diff": d["fd_check"]["max_rel_diff"], "fd_directions": 2,
                          "workload": d["config"]["workload"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
