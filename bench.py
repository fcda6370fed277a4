"""Benchmark: MLMG solve to 1e-10 relative residual on the 512^3
variable-coefficient ABecLaplacian (homogeneous Dirichlet, max_grid_size 128,
64 boxes), mixed precision (FP32 V-cycle in FP64 defect correction) against
plain FP64, plus smoother GB/s.

One step = one full solve from a zero initial guess with inputs resident in
HBM.  Prints one JSON line.  Single GPU only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLMG solve s to 1e-10 resid, 512^3 ABec; mixed vs FP64 speedup; smoother GB/s"
HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def face_count(boxes):
    nf = 0
    for b in boxes:
        n = [b[3 + d] - b[d] + 1 for d in range(3)]
        nf += (n[0] + 1) * n[1] * n[2] + n[0] * (n[1] + 1) * n[2] + n[0] * n[1] * (n[2] + 1)
    return nf


def host_cores():
    """(logical CPUs this process may run on, physical cores among them): the
    affinity mask, and the threads-per-core ratio of /proc/cpuinfo ("siblings"
    over "cpu cores" of a package)."""
    logical = len(os.sched_getaffinity(0))
    siblings = cores = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "siblings" and siblings is None:
                    siblings = int(v)
                elif k == "cpu cores" and cores is None:
                    cores = int(v)
    except OSError:
        pass
    per_core = max(1, siblings // cores) if siblings and cores else 1
    return logical, max(1, logical // per_core)


def cpu_baseline(cfg, inputs):
    """One complete FP64 solve of the same workload by the CPU oracle (a C++/
    OpenMP restatement of the reference's CPU MLMG, agglomeration 8 as the
    reference uses on CPUs), from a zero initial guess to the same tolerance;
    setup (allocation, coefficient coarsening) and iterations timed apart."""
    from oracle import pyoracle as po
    a, bx, by, bz, rhs = inputs
    prm = po.make_params(cfg.n, op=cfg.op, max_grid_size=cfg.max_grid_size, agg_grid_size=8,
                         maxorder=cfg.maxorder, alpha=cfg.alpha, beta=cfg.beta)
    t = time.perf_counter()
    _, st = po.solve(prm, rhs, a, bx, by, bz)
    wall = time.perf_counter() - t
    logical, physical = host_cores()
    threads = int(os.environ.get("OMP_NUM_THREADS", logical))
    return {"value": round(st["solve_seconds"], 3), "unit": "s", "cores": threads, "kind": "port",
            "host_logical_cpus": logical, "host_physical_cores": physical,
            "setup_s": round(st["setup_seconds"], 3), "iterations_s": round(st["iter_seconds"], 3),
            "iters": st["niters"], "wall_s": round(wall, 3),
            "sample": (f"one complete CPU-oracle FP64 solve of the {cfg.n[0]}^3 workload to rel tol {cfg.tol_rel} "
                       f"({st['niters']} V-cycle iterations), {threads} OpenMP threads on {logical} logical CPUs "
                       f"({physical} physical cores): setup {st['setup_seconds']:.2f} s + iterations "
                       f"{st['iter_seconds']:.2f} s")}


def hbm_probe(nbytes=1 << 31, reps=10):
    """Practical HBM ceiling: hand-written streaming kernels (16-byte
    non-temporal accesses) over nbytes: copy (read + write), read only, write
    only, each in GB/s of bytes moved."""
    import ctypes
    import torch
    from amrex_amd import capi
    lib = capi.load()
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    out = {}
    for mode, name, moved in ((0, "copy", 2 * nbytes), (1, "read", nbytes), (2, "write", nbytes)):
        ms = ctypes.c_double()
        lib.mgx_hbm_probe(mode, ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(dst.data_ptr()), nbytes, reps,
                          ctypes.byref(ms))
        capi.check("hbm probe: ")
        out[name] = round(moved / (ms.value * 1e-3) / 1e9, 1)
    del src, dst
    torch.cuda.empty_cache()
    return out


def sweep_traffic(prec):
    """HBM bytes per launch of the finest-level fused sweep of the 512^3
    workload, from the FETCH_SIZE / WRITE_SIZE counter passes recorded in
    profiles/gsrb_sweep_abec512.md (scaled by the calibration pass there)."""
    import re
    path = os.path.join(ROOT, "profiles", "gsrb_sweep_abec512.md")
    try:
        text = open(path).read()
    except OSError:
        return None
    m = re.search(r"HBM traffic per launch:.*?" + prec + r" (\d+) B", text)
    return int(m.group(1)) if m else None


def smoother_rates(prob, cfg, sweeps):
    """Finest-level smoother sweep time (HIP events on the library stream) and
    algorithmic bandwidth: s * (4 N_c + N_f) bytes per red+black sweep (phi
    read + write, rhs, a, and the three face coefficients).  N_f counts the
    faces of the storage: the domain's on the single-FAB (fused) layout, each
    box's own on the box layout."""
    import numpy as np
    from amrex_amd import mlmg
    boxes = prob.linop.levelInfo(0)[1]
    nc = int(np.prod(cfg.n))
    n = cfg.n
    sm = {}
    for prec, s in (("double", 8), ("mixed", 4)):
        ms_sweep, ms_gsrb, ms_fill = mlmg.time_smoother(prob.linop, 0, prec, sweeps)
        fused = ms_fill == 0.0
        nf = face_count([(0, 0, 0, n[0] - 1, n[1] - 1, n[2] - 1)] if fused else boxes)
        bytes_sweep = s * (4 * nc + nf)
        sm[prec] = {"ms_sweep": ms_sweep, "ms_gsrb_launch": ms_gsrb, "ms_fill_launch": ms_fill,
                    "fused": fused, "bytes": bytes_sweep, "gbps": bytes_sweep / (ms_sweep * 1e-3) / 1e9}
        log(f"[bench] smoother {prec}: {ms_sweep:.3f} ms/sweep ({sm[prec]['gbps']:.0f} GB/s algorithmic), "
            f"colour launch {ms_gsrb:.3f} ms, ghost fill {ms_fill:.3f} ms")
        # the two-sweep launch the V-cycle uses for consecutive FP32 sweeps, when the level allows it
        ms_pair = None
        if prec == "mixed":
            try:
                ms_pair = mlmg.time_sweep_pairs(prob.linop, 0, prec, max(2, sweeps // 2))
            except Exception:
                ms_pair = None
        if ms_pair:
            sm[prec]["ms_pair_launch"] = ms_pair
            sm[prec]["gbps_pair"] = 2 * bytes_sweep / (ms_pair * 1e-3) / 1e9
            log(f"[bench] smoother {prec}: two-sweep launch {ms_pair:.3f} ms ({ms_pair / 2:.3f} ms per sweep, "
                f"{sm[prec]['gbps_pair']:.0f} GB/s algorithmic over both sweeps)")
    return sm, boxes


def roofline_line(sm, fused, args):
    """Roofline of the dominant kernel, the finest-level FP32 smoother: the
    two-sweep launch where the V-cycle uses it (algorithmic bytes of both
    sweeps, each s * (4 N_c + N_f)), else the single-sweep launch."""
    m = sm["mixed"]
    full = fused and args.config == "abec512" and not args.n
    if "gbps_pair" in m:
        return {"bound": "hbm",
                "kernel": "FP32 two-sweep red+black GSRB march on the finest level (k_gsrb_sweep2<float>)",
                "achieved": round(m["gbps_pair"], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(m["gbps_pair"] / HBM_PEAK_GBPS, 4),
                "bytes_per_launch": 2 * m["bytes"], "ms_per_launch": round(m["ms_pair_launch"], 4),
                "traffic": sweep_traffic("fp32 pair") if full else None,
                "traffic_source": ("FETCH_SIZE + WRITE_SIZE counter passes over this kernel, recorded in "
                                   "profiles/gsrb_sweep_abec512.md (not counted in this run)")}
    return {"bound": "hbm",
            "kernel": ("FP32 single-pass red+black GSRB sweep on the finest level (k_gsrb_sweep<float>)"
                       if fused else "FP32 GSRB smoother sweep on the finest level (2 ghost fills + red + black)"),
            "achieved": round(m["gbps"], 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(m["gbps"] / HBM_PEAK_GBPS, 4),
            "bytes_per_launch": m["bytes"], "ms_per_launch": round(m["ms_sweep"], 4),
            "traffic": sweep_traffic("fp32") if full else None,
            "traffic_source": ("FETCH_SIZE + WRITE_SIZE counter passes over this kernel, recorded in "
                               "profiles/gsrb_sweep_abec512.md (not counted in this run)")}


def kernel_only(prob, cfg, args):
    import torch
    sm, _ = smoother_rates(prob, cfg, args.smoother_sweeps)
    # elementwise pass over 2^28 doubles (2 GiB read, 2 GiB written), for
    # scaling byte counters (a kernel of its own name: dst = src * 1)
    src = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 5
    del src, dst
    out = {"smoother": sm, "copy_bytes_each_way": (1 << 28) * 8, "copy_ms": ms,
           "copy_gbps_read_plus_write": 2 * (1 << 28) * 8 / (ms * 1e-3) / 1e9, "hbm_probe_gbps": hbm_probe()}
    print(json.dumps(out), flush=True)


def bench_amr(args):
    """Two-level AMR composite Poisson: coarse 256^3 periodic (max_grid_size 64),
    40 seeded fine boxes (ratio 2), Gaussian-blob rhs; FP64 and mixed solves to
    1e-10.  One JSON line with the solve seconds of both forms."""
    import numpy as np
    import torch
    from amrex_amd import mlmg, workload as wl
    nc = args.n or 256
    boxes = wl.amr_boxes(nc, 40 if nc >= 256 else 5)
    r0, r1 = wl.amr_rhs(nc, boxes)
    # inputs resident in device memory before the timed region
    d0 = torch.as_tensor(r0, device="cuda")
    d1 = [torch.as_tensor(b, device="cuda") for b in r1]
    out = {}
    for prec in ("double", "mixed"):
        prob = mlmg.AmrProblem(nc, 64, boxes, precision=prec)
        times = []
        for it in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            prob.solve_device(d0, d1)
            torch.cuda.synchronize()
            if it >= args.warmup:
                times.append(time.perf_counter() - t)
        st = prob.mlmg.stats()
        out[prec] = {"median_s": float(np.median(times)), "iters": st["niters"],
                     "iter_s": st["iter_seconds"]}
        log(f"[bench amr] {prec}: {['%.4f' % t for t in times]} s (device tensors in and out), {st['niters']} "
            f"iterations, iteration loop {st['iter_seconds']:.4f} s")
    fine_cells = sum((b[3] - b[0] + 1) * (b[4] - b[1] + 1) * (b[5] - b[2] + 1) for b in boxes)
    print(json.dumps({"metric": "two-level AMR composite Poisson solve s to 1e-10, coarse 256^3 + 40 fine boxes",
                      "value": round(out["mixed"]["median_s"], 5), "unit": "s", "n_gpus": 1, "higher_is_better": False,
                      "dtype": "f32+f64", "data": "synthetic",
                      "config": {"workload": f"amr: coarse {nc}^3 periodic Poisson, {len(boxes)} fine boxes "
                                             f"({fine_cells} fine cells), 3 Gaussian blobs"},
                      "fp64": out["double"], "mixed": out["mixed"]}), flush=True)


def bench_adjoint(args):
    """Adjoint gradients: 256^3 ABecLaplacian (homogeneous Dirichlet, max_grid_size
    64), batch of 8 seeded rhs, J = 0.5|phi - phi_t|^2; time of the forward
    batch + adjoint batch + coefficient gradients (mixed precision)."""
    import numpy as np
    import torch
    from amrex_amd import adjoint, mlmg, workload as wl
    cfg = wl.CONFIGS["abec256"]
    a, bx, by, bz, _ = wl.make_inputs(cfg)
    rhs, tgt = wl.adjoint_inputs(cfg.n, 8)
    dev = torch.device("cuda")
    prob = mlmg.Problem(cfg.n, cfg.max_grid_size, op="abec", a=a, bx=bx, by=by, bz=bz, precision="mixed")
    tb = [torch.as_tensor(c, device=dev) for c in (bx, by, bz)]
    ta, tr, tt = (torch.as_tensor(x, device=dev) for x in (a, rhs, tgt))
    times = []
    for it in range(args.warmup + args.steps):
        for t in tb:
            t.requires_grad_(True)
            t.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        phi = adjoint.solve(prob, tr, ta, *tb, tol_rel=1e-10)
        J = 0.5 * ((phi - tt) ** 2).sum()
        J.backward()
        torch.cuda.synchronize()
        if it >= args.warmup:
            times.append(time.perf_counter() - t0)
    log(f"[bench adjoint] forward+adjoint batch of 8 at 256^3: {['%.4f' % t for t in times]} s")
    print(json.dumps({"metric": "adjoint gradient s (8 forward + 8 adjoint MLMG solves + gradients), 256^3 ABec",
                      "value": round(float(np.median(times)), 5), "unit": "s", "n_gpus": 1, "higher_is_better": False,
                      "dtype": "f32+f64", "data": "synthetic",
                      "config": {"workload": "adjoint256: abec256 operator, 8 rhs (seeds 501-508), phi_t seed 599, "
                                             "mixed precision, rel tol 1e-10"}}), flush=True)


def bench_variants(args, kind):
    """The abec512 workload solved by the other drivers of the same V-cycle: kind
    "fmg" = MLMG with one F-cycle iteration (setMaxFmgIter(1)) then V-cycles;
    kind "gmres" = GMRESMLMG (restart 30, one V-cycle per preconditioner
    application; mixed precision: FP32 V-cycles, flexible GMRES).  FP64 and
    mixed, each against the plain V-cycle solve of the same precision."""
    import numpy as np
    import torch
    from amrex_amd import mlmg, workload as wl
    cfg = wl.CONFIGS["abec512"]
    if args.n:
        cfg = wl.Config(**{**cfg.__dict__, "n": (args.n,) * 3})
    a, bx, by, bz, rhs = wl.make_inputs(cfg)
    out = {}
    for prec in ("double", "mixed"):
        prob = mlmg.Problem(cfg.n, cfg.max_grid_size, op=cfg.op, a=a, bx=bx, by=by, bz=bz, alpha=cfg.alpha,
                            beta=cfg.beta, maxorder=cfg.maxorder, precision=prec)
        prob.rhs.fromGlobal(torch.as_tensor(rhs, device="cuda").contiguous(), cfg.n)
        if kind == "gmres":
            gm = mlmg.GMRESMLMG(prob.mlmg)

        def run(variant):
            prob.sol.setVal(0.0)
            torch.cuda.synchronize()
            t = time.perf_counter()
            if variant == "plain":
                prob.mlmg.setMaxFmgIter(0)
                prob.mlmg.solve(prob.sol, prob.rhs, cfg.tol_rel, 0.0)
            elif variant == "fmg":
                prob.mlmg.setMaxFmgIter(1)
                prob.mlmg.solve(prob.sol, prob.rhs, cfg.tol_rel, 0.0)
            else:
                gm.solve(prob.sol, prob.rhs, cfg.tol_rel, 0.0)
            torch.cuda.synchronize()
            return time.perf_counter() - t

        for variant in ("plain", kind):
            for _ in range(args.warmup):
                run(variant)
            times = [run(variant) for _ in range(args.steps)]
            st = gm.stats() if variant == "gmres" else prob.mlmg.stats()
            iters = st["niters"]
            out[f"{variant}_{prec}"] = {"median_s": round(float(np.median(times)), 5), "iters": iters}
            log(f"[bench {kind}] {variant} {prec}: {['%.4f' % t for t in times]} s, {iters} iterations")
        prob.mlmg.setMaxFmgIter(0)
    what = ("GMRESMLMG (right-preconditioned GMRES(30), V-cycle preconditioner)" if kind == "gmres"
            else "MLMG with one F-cycle iteration")
    print(json.dumps({"metric": f"{what} solve s to rel tol {cfg.tol_rel}, abec512",
                      "value": out[f"{kind}_mixed"]["median_s"], "unit": "s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "dtype": "f32+f64",
                      "data": "synthetic",
                      "config": {"workload": f"{cfg.name}: {cfg.n[0]}^3 variable-coefficient ABecLaplacian, "
                                             f"max_grid_size {cfg.max_grid_size}, seed {cfg.seed}"},
                      "results": out}), flush=True)


def bench_io(args):
    """VisMF::Write / VisMF::Read and WriteSingleLevelPlotfile of the abec512
    solution MultiFab (512^3 FP64, boxes of max_grid_size, one ghost layer),
    beside the two bounds of the path: a pinned device-to-host copy of the
    same bytes and a plain file write (and read) of them from (into) pinned
    host memory."""
    import shutil
    import tempfile

    import numpy as np
    import torch
    from amrex_amd import mlmg, workload as wl
    cfg = wl.CONFIGS["abec512"]
    n = (args.n,) * 3 if args.n else cfg.n
    ba = mlmg.BoxArray((0, 0, 0), [v - 1 for v in n]).maxSize(cfg.max_grid_size)
    dm = mlmg.DistributionMapping(ba)
    mf = mlmg.MultiFab(ba, dm, 1, 1)
    g = torch.Generator(device="cuda").manual_seed(cfg.seed)
    mf.fromGlobal(torch.randn((n[2], n[1], n[0]), dtype=torch.float64, device="cuda", generator=g), n)
    geom = mlmg.Geometry(n)
    nbytes = 8 * n[0] * n[1] * n[2]
    ref_sum, ref_max = mf.sum(), mf.norminf()
    root = tempfile.mkdtemp(prefix="mgx_io_", dir=os.environ.get("TMPDIR", "/tmp"))
    times = {"vismf_write": [], "vismf_read": [], "plotfile_write": [], "d2h_pinned": [], "file_write": [],
             "file_read": []}
    try:
        host = torch.empty(nbytes // 8, dtype=torch.float64, pin_memory=True)
        dev = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda")
        for it in range(args.warmup + args.steps):
            rec = it >= args.warmup

            def timed(key, fn):
                torch.cuda.synchronize()
                t = time.perf_counter()
                fn()
                torch.cuda.synchronize()
                if rec:
                    times[key].append(time.perf_counter() - t)

            name = os.path.join(root, "sol%d" % it)
            timed("vismf_write", lambda: mf.write(name))
            box = {}
            timed("vismf_read", lambda: box.setdefault("mf", mlmg.MultiFab.read(name)))
            back = box.pop("mf")
            assert back.sum() == ref_sum and back.norminf() == ref_max, "read-back differs"
            del back
            timed("plotfile_write", lambda: mlmg.WriteSingleLevelPlotfile(os.path.join(root, "plt%d" % it), mf,
                                                                          ["phi"], geom, 0.0, it))
            timed("d2h_pinned", lambda: host.copy_(dev))
            raw = os.path.join(root, "raw%d" % it)

            def plain():
                with open(raw, "wb") as f:
                    f.write(host.numpy().data)
            timed("file_write", plain)

            def plain_read():
                with open(raw, "rb", buffering=0) as f:
                    f.readinto(host.numpy().data.cast("B"))
            timed("file_read", plain_read)
            for p in (name + "_H", name + "_D_00000", raw):
                os.remove(p)
            shutil.rmtree(os.path.join(root, "plt%d" % it))
            log(f"[bench io] step {it}: " + ", ".join(f"{k} {v[-1]:.3f} s" for k, v in times.items() if rec and v))
    finally:
        shutil.rmtree(root, ignore_errors=True)
    rate = {k: round(nbytes / float(np.median(v)) / 1e9, 2) for k, v in times.items()}
    print(json.dumps({"metric": "VisMF write of a 512^3 FP64 MultiFab from HBM to a file", "value": rate["vismf_write"],
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "higher_is_better": True, "dtype": "f64", "data": "synthetic",
                      "config": {"workload": f"{n[0]}^3 one-component MultiFab, one ghost layer, "
                                             f"max_grid_size {cfg.max_grid_size}, files under {root}"},
                      "rates_GBps": rate, "bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="abec512",
                    help="abec512 (headline), abec256, poisson64, amr256 (two-level composite), adjoint256, "
                         "fmg512 / gmres512 (the abec512 workload by the F-cycle / GMRES drivers), "
                         "io512 (VisMF and plotfile write/read of the 512^3 solution MultiFab)")
    ap.add_argument("--n", type=int, default=0, help="override the cubic grid size (testing only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--smoother-sweeps", type=int, default=10)
    ap.add_argument("--kernel-only", action="store_true",
                    help="no solves: time the finest-level smoother sweeps (FP64, FP32) and a device copy "
                         "of known size (the byte-counter calibration of a profiler run)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU; there is no multi-GPU path (the solve would shard by boxes "
              "across GPUs).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from amrex_amd import mlmg, workload as wl

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.config == "amr256":
        bench_amr(args)
        return
    if args.config == "adjoint256":
        bench_adjoint(args)
        return
    if args.config == "io512":
        bench_io(args)
        return
    if args.config in ("fmg512", "gmres512"):
        bench_variants(args, args.config[:-3])
        return
    cfg = wl.CONFIGS[args.config]
    if args.n:
        cfg = wl.Config(**{**cfg.__dict__, "n": (args.n,) * 3})
    t0 = time.time()
    inputs = wl.make_inputs(cfg)
    a, bx, by, bz, rhs = inputs
    log(f"[bench] inputs {cfg.n} generated in {time.time() - t0:.1f} s")
    prob = mlmg.Problem(cfg.n, cfg.max_grid_size, op=cfg.op, a=a, bx=bx, by=by, bz=bz,
                        alpha=cfg.alpha, beta=cfg.beta, maxorder=cfg.maxorder)
    rhs_t = torch.as_tensor(rhs, device="cuda").contiguous()
    prob.rhs.fromGlobal(rhs_t, cfg.n)
    log(f"[bench] operator ready, {prob.linop.NMGLevels()} MG levels")
    if args.kernel_only:
        kernel_only(prob, cfg, args)
        return

    def one_solve():
        prob.sol.setVal(0.0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        prob.mlmg.solve(prob.sol, prob.rhs, cfg.tol_rel, 0.0)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    # FP64 (the reference's iteration), mixed with one FP32 V-cycle per FP64
    # refinement step, and mixed with one or two per step (the headline form)
    res = {}
    sols = {}
    for key, prec, refine in (("double", "double", 1), ("mixed1", "mixed", 1), ("mixed", "mixed", 0)):
        prob.mlmg.setPrecision(prec)
        prob.mlmg.setRefinementVcycles(refine)
        for _ in range(args.warmup):
            one_solve()
        if key == "mixed":
            torch.cuda.synchronize()
            t_start = time.perf_counter()
        times = [one_solve() for _ in range(args.steps)]
        if key == "mixed":
            torch.cuda.synchronize()
            total = time.perf_counter() - t_start
        st = prob.mlmg.stats()
        out = torch.zeros((cfg.n[2], cfg.n[1], cfg.n[0]), dtype=torch.float64, device="cuda")
        prob.sol.toGlobal(out, cfg.n)
        sols[key] = out
        max_norm = max(st["rhsnorm"], st["init_resnorm"])
        res[key] = {"times": times, "median": float(np.median(times)), "iters": st["niters"],
                    "vcycles": st["nvcycles"], "final_rel_resid": st["final_resnorm"] / max_norm}
        log(f"[bench] {key}: solves {['%.4f' % t for t in times]} s, {st['niters']} steps, "
            f"{st['nvcycles']} V-cycles, rel resid {res[key]['final_rel_resid']:.3e}")
    prob.mlmg.setRefinementVcycles(1)
    err = float(torch.linalg.norm(sols["mixed"] - sols["double"]) / torch.linalg.norm(sols["double"]))
    err1 = float(torch.linalg.norm(sols["mixed1"] - sols["double"]) / torch.linalg.norm(sols["double"]))
    del sols

    sm, boxes = smoother_rates(prob, cfg, args.smoother_sweeps)
    fused = sm["mixed"]["fused"]

    mixed = res["mixed"]
    line = {
        "metric": METRIC if cfg.name == "abec512" else METRIC.replace("512^3 ABec", cfg.name),
        "value": round(mixed["median"], 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"{cfg.name}: {cfg.n[0]}x{cfg.n[1]}x{cfg.n[2]} "
                                + ("variable-coefficient ABecLaplacian, homogeneous Dirichlet"
                                   if cfg.op == "abec" else "periodic Poisson, 8 Fourier modes")
                                + f", maxorder {cfg.maxorder}, "
                                f"max_grid_size {cfg.max_grid_size} ({len(boxes)} boxes), seed {cfg.seed}, "
                                f"rel tol {cfg.tol_rel}, mixed precision (FP32 V-cycles, FP64 refinement, "
                                f"one or two V-cycles per refinement step)"),
                   "mg_levels": prob.linop.NMGLevels()},
        "fp64_solve_s": round(res["double"]["median"], 5),
        "mixed_speedup_vs_fp64": round(res["double"]["median"] / mixed["median"], 3),
        "mixed_one_vcycle_per_step_s": round(res["mixed1"]["median"], 5),
        "mixed_one_vcycle_per_step_speedup_vs_fp64": round(res["double"]["median"] / res["mixed1"]["median"], 3),
        "iters_fp64": res["double"]["iters"],
        "iters_mixed": mixed["iters"],
        "vcycles_mixed": mixed["vcycles"],
        "iters_mixed_one_vcycle_per_step": res["mixed1"]["iters"],
        "vcycles_per_s_mixed": round(mixed["vcycles"] / mixed["median"], 2),
        "vcycles_per_s_fp64": round(res["double"]["iters"] / res["double"]["median"], 2),
        "final_rel_resid_mixed": mixed["final_rel_resid"],
        "final_rel_resid_fp64": res["double"]["final_rel_resid"],
        "mixed_vs_fp64_rel_l2": err,
        "mixed_one_vcycle_per_step_vs_fp64_rel_l2": err1,
        "smoother_gbps_fp64": round(sm["double"]["gbps"], 1),
        "smoother_gbps_fp32": round(sm["mixed"].get("gbps_pair", sm["mixed"]["gbps"]), 1),
        "smoother_gbps_fp32_single_sweep_launch": round(sm["mixed"]["gbps"], 1),
        "smoother_ms_per_sweep_fp32": round(sm["mixed"].get("ms_pair_launch", 2 * sm["mixed"]["ms_sweep"]) / 2, 4),
        "smoother_ms_per_sweep_fp64": round(sm["double"]["ms_sweep"], 4),
        "roofline": roofline_line(sm, fused, args),
        "hbm_probe_gbps": hbm_probe(),
    }
    if not args.no_cpu:
        log("[bench] CPU baseline sample ...")
        line["cpu_baseline"] = cpu_baseline(cfg, inputs)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
