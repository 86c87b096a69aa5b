#!/usr/bin/env python3
"""Benchmark: mixed-precision VPGMRES solve time on one MI355X.

Workload (the configuration the headline metric is quoted on): 3-D 7-point
Poisson on a 160^3 grid (n = 4,096,000, nnz = 28,518,400), b ~ U(-1,1)
(seed 3), x0 = 0; outer FP64 flexible GMRES(m=20) (-R 21), inner 10 steps of
FP32 GMRES on an FP32 copy of A (-s gmres -l 11 -r 12 -e 0), stop at 1e-10
relative residual.  One step = one full solve from x0 = 0 with inputs
resident in HBM.

Reported with it (outside the timed region): plain FP64 GMRES(20) on the
same GPU (speed-up), the relative error of the mixed solution against an FP64
GMRES solution run to 1e-12, SpMV bandwidth, the roofline of the dominant
kernel (hipEvents around its launches inside the timed solves), and the CPU
oracle timed on a bounded sample on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# dominant kernel of the headline solve (the inner FP32 GMRES step's SpMV
# with the Gram-Schmidt projection fused behind it, reading int16 column
# offsets: the stencil's columns lie within 25,600 of their rows), and the committed counter
# summary (two rocprofv3 --pmc runs, copy-calibrated) that gives its HBM traffic
DOM_KERNEL = "ks::k_spmv_stream<float, 1, 0, ks::EpiDivProj<float>, short>"
TRAFFIC_FILE = os.path.join("profiles", "k_spmv_stream_fp32_poisson160_traffic.txt")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (None when the summary is absent or was made for another kernel)."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel.replace(" ", ""):
            return float(f[2])
    return None
METRIC = "solve s to 1e-10 rel-res, mixed vs FP64 on GPU; SpMV HBM GB/s; adjoint grad s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=160, help="Poisson grid edge (160 = the benchmark)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 GMRES comparison")
    ap.add_argument("--cpu-steps", type=int, default=0,
                    help="outer steps in the CPU sample (0: the whole solve, about 10 s on 16 cores)")
    ap.add_argument("--cpu-runs", type=int, default=3, help="CPU baseline runs (the median is reported)")
    ap.add_argument("--prof-steps", type=int, default=3, help="profiled solves for the roofline")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the per-configuration SpMV and adjoint timings")
    ap.add_argument("--suite", action="store_true",
                    help="also run the per-configuration SpMV, parity-solve and adjoint measurements "
                         "(one JSON line each, after the headline line)")
    ap.add_argument("--fd-dirs", type=int, default=64, help="finite-difference directions per adjoint case")
    ap.add_argument("--spmv-config", default=None,
                    help="only the SpMV measurement of one configuration (its default layout, A and A^T), "
                         "one JSON line, no headline: e.g. zipf, for a profiler run")
    ap.add_argument("--spmv-layouts", default="",
                    help="with --spmv-config: also time these layouts (comma-separated names of "
                         "solver_cpp_amd.api.Matrix.LAYOUTS, e.g. panels_merge,panels_wg,panels_wave)")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint timings (forward + gradient per case), one JSON line, no headline: "
                         "for a profiler run")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="MiB of one plain device copy at the end (counter calibration)")
    return ap.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch
    from solver_cpp_amd import api, gen
    from solver_cpp_amd._abi import KERNEL_CLASSES

    if a.adjoint_only:
        ctx = api.Context(0)
        print(json.dumps({"suite": "adjoint", "adjoint": adjoint_times(ctx)}), flush=True)
        ctx.close()
        return
    if a.spmv_config:
        P = gen.problem(a.spmv_config)
        n, nnz = len(P["ptr"]) - 1, len(P["col"])
        ctx = api.Context(0)
        A = api.Matrix(ctx, P["ptr"], P["col"], P["val"], fp32=True, transpose=True)
        print(json.dumps({"suite": "spmv", "config": a.spmv_config, "n": n, "nnz": nnz,
                          "spmv_kind": A.info()["spmv_kind"], "fp64": spmv_gbps(A, n, nnz, 64),
                          "fp32": spmv_gbps(A, n, nnz, 32), "fp64_transpose": spmv_gbps(A, n, nnz, 64, trans=True)}),
              flush=True)
        A.close()
        for lay in filter(None, a.spmv_layouts.split(",")):
            t = time.perf_counter()
            A = api.Matrix(ctx, P["ptr"], P["col"], P["val"], fp32=True, layout=lay)
            print(json.dumps({"suite": "spmv_layout", "config": a.spmv_config, "layout": lay,
                              "setup_s": round(time.perf_counter() - t, 2), "spmv_kind": A.info()["spmv_kind"],
                              "fp64": spmv_gbps(A, n, nnz, 64), "fp32": spmv_gbps(A, n, nnz, 32)}), flush=True)
            A.close()
        calib_copy(a)
        ctx.close()
        return

    g = a.grid
    P = gen.problem("poisson3d", scale=g)
    n, nnz = len(P["ptr"]) - 1, len(P["col"])
    ctx = api.Context(0)
    A = api.Matrix(ctx, P["ptr"], P["col"], P["val"], fp32=True)
    b = torch.from_numpy(P["b"]).cuda()
    mixed = dict(outer_restart=21, outer_maxloop=5000, outer_eps=1e-10, inner_maxloop=11, inner_restart=12,
                 inner_eps=0.0, inner_precision=32)
    dom = KERNEL_CLASSES.index("spmv32")

    def run(outer, inner="NO", profile=0, **kw):
        return api.solve(ctx, A, b, outer, inner, history=False, profile=profile, **kw)

    def wall(outer, inner="NO", **kw):
        """seconds of one solve, timed as the headline is (device-synchronised
        wall clock around the call)"""
        torch.cuda.synchronize()
        t = time.perf_counter()
        x, r, _ = run(outer, inner, **kw)
        torch.cuda.synchronize()
        return time.perf_counter() - t, x, r

    for _ in range(a.warmup):
        run("vpgmres", "gmres", **mixed)
    # the timed region: `steps` unprofiled solves
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        x, r0, _ = run("vpgmres", "gmres", **mixed)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / a.steps
    # roofline of the dominant kernel from a separate pass of profiled solves:
    # every launch of it carries a kernel-timestamp event pair.  Its mean
    # algorithmic bytes per launch by the byte model of the metric (int32
    # CSR): SpMV 8*nnz + 12*n + 4, plus 4*n*nv for the nv basis vectors the
    # fused projection reads (nv = k + 1 at inner step k).  The library
    # counts the bytes of the format the kernel reads, (4 + ib)*nnz + ... with
    # ib bytes per column index (2: int16 offsets, 4: int32 columns); that
    # rate, the kernel's use of HBM bandwidth, is reported beside it.
    prof = [run("vpgmres", "gmres", profile=1 << dom, **mixed)[1] for _ in range(a.prof_steps)]
    cls = KERNEL_CLASSES[dom]
    ms_tot = sum(rr["prof"].get(cls, {}).get("ms", 0.0) for rr in prof)
    launches = sum(rr["prof"].get(cls, {}).get("launches", 0) for rr in prof)
    by_tot = sum(rr["prof"].get(cls, {}).get("bytes", 0.0) for rr in prof)
    bytes_per = by_tot / max(launches, 1)
    ms_per = ms_tot / max(launches, 1)
    ib = A.index_bytes(False, 32)
    bytes_per32 = bytes_per + (4 - ib) * nnz  # the same launches, int32 column indices
    achieved = bytes_per32 / (ms_per * 1e-3) / 1e9 if launches else None
    moved = bytes_per / (ms_per * 1e-3) / 1e9 if launches else None

    out = {
        "metric": METRIC, "value": round(sec, 6), "unit": "s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(sec * 1e3, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 outer / f32 inner", "data": "synthetic",
        "config": {"workload": f"VPGMRES(20)+10xFP32 GMRES, 3-D 7-pt Poisson {g}^3, tol 1e-10",
                   "n": n, "nnz": nnz, "outer_steps": r0["steps"], "inner_steps": r0["inner_steps"],
                   "iters": r0["iters"], "relres_true": r0["relres_true"]},
    }
    tr = measured_traffic(DOM_KERNEL)
    out["roofline"] = {"bound": "hbm", "kernel": DOM_KERNEL,
                       "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                       "traffic": tr, "traffic_source": TRAFFIC_FILE if tr else None,
                       "bytes_per_launch": bytes_per32, "ms_per_launch": ms_per, "launches": launches,
                       "byte_model": "int32 CSR (8*nnz + 12*n + 4 + 4*n*nv)",
                       "index_bytes": ib,
                       "format": {"what": "bytes of the format the kernel reads (16-bit column indices when "
                                          "index_bytes is 2): its use of HBM bandwidth",
                                  "bytes_per_launch": bytes_per, "GBps": round(moved, 1) if moved else None,
                                  "frac": round(moved / HBM_PEAK_GBS, 4) if moved else None},
                       "profiled_solves": a.prof_steps,
                       "profiled_s_per_solve": round(sum(rr["seconds"] for rr in prof) / max(len(prof), 1), 6)}
    out["solve_bytes_GBps"] = round(r0["bytes"] / r0["seconds"] / 1e9, 1)

    if not a.no_fp64:
        # FP64 GMRES(20) on the same GPU (median of 3, timed as the headline),
        # the FP64-inner VPGMRES (the reference's own precision: what FP32 in
        # the inner solve alone buys), and the FP64 solution run to 1e-12
        t64 = [wall("gmres", outer_restart=21, outer_maxloop=200000, outer_eps=1e-10) for _ in range(3)]
        s64 = sorted(t[0] for t in t64)[1]
        r64 = t64[0][2]
        f64in = dict(mixed, inner_precision=64)
        tin = [wall("vpgmres", "gmres", **f64in) for _ in range(3)]
        sin = sorted(t[0] for t in tin)[1]
        rin = tin[0][2]
        xo, ro, _ = run("gmres", outer_restart=21, outer_maxloop=400000, outer_eps=1e-12)
        out["fp64_gmres"] = {"seconds": round(s64, 4), "iters": r64["iters"], "relres_true": r64["relres_true"],
                             "runs": 3, "stat": "median"}
        out["speedup_vs_fp64_gmres"] = round(s64 / sec, 3)
        out["vpgmres_fp64_inner"] = {"seconds": round(sin, 4), "outer_steps": rin["steps"],
                                     "inner_steps": rin["inner_steps"], "relres_true": rin["relres_true"],
                                     "runs": 3, "stat": "median", "speedup_of_fp32_inner": round(sin / sec, 3)}
        out["rel_error_vs_fp64_oracle"] = float(torch.linalg.norm(x - xo) / torch.linalg.norm(xo))
    out["spmv"] = {"poisson3d_160": spmv_line(A, n, nnz, trans=False)}
    A.close()
    if not a.headline_only:
        out["hbm_ceiling"] = hbm_ceiling(ctx, achieved, moved)
        out["spmv"].update(spmv_configs(ctx))
        out["adjoint"] = adjoint_times(ctx)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, r0, a)
    if a.calib_copy:
        out["calib_copy_bytes"] = calib_copy(a)
    print(json.dumps(out), flush=True)
    if a.suite:
        suite(ctx, a)


def spmv_line(A, n, nnz, trans):
    """SpMV GB/s of one matrix: FP64 and FP32, and A^T in FP64 when built."""
    rec = {"fp64": spmv_gbps(A, n, nnz, 64), "fp32": spmv_gbps(A, n, nnz, 32)}
    if trans:
        rec["fp64_transpose"] = spmv_gbps(A, n, nnz, 64, trans=True)
    return rec


_ZIPF = {}


def zipf_problem():
    """The 4,000,000-row Zipf matrix, generated once per run."""
    from solver_cpp_amd import gen
    if not _ZIPF:
        _ZIPF.update(gen.problem("zipf"))
    return _ZIPF


def spmv_configs(ctx):
    """SpMV bandwidth on the other benchmark matrices, each on the layout the
    library picks for it: 3-D convection-diffusion 128^3 (Pe 0.5) and 192^3
    (Pe 0.8), and the 4,000,000-row Zipf matrix (column panels)."""
    from solver_cpp_amd import api, gen
    out = {}
    for name, cfg in (("convdiff3d_128", "convdiff3d"), ("convdiff3d_pe08_192", "convdiff3d_pe08"),
                      ("zipf_4m", "zipf")):
        P = zipf_problem() if cfg == "zipf" else gen.problem(cfg)
        n, nnz = len(P["ptr"]) - 1, len(P["col"])
        A = api.Matrix(ctx, P["ptr"], P["col"], P["val"], fp32=True, transpose=True)
        out[name] = dict(spmv_line(A, n, nnz, trans=True), layout=A.info()["spmv_kind"])
        A.close()
    return out


def adjoint_times(ctx):
    """Forward solve and gradient (adjoint solve + SDDMM) seconds:
    - 1024^2 anisotropic diffusion (SPD): VPCG with FP32 inner CG, adjoint CG;
    - 192^3 convection-diffusion Pe 0.8: mixed VPGMRES(20), adjoint GMRES(20)
      on the explicit transpose;
    - the Zipf matrix, row-equilibrated (D^-1 A): the same pair of solvers.
    All solves to 1e-14 (the accuracy the finite-difference checks need);
    each case timed after one untimed forward + backward."""
    import torch
    from solver_cpp_amd import autograd as AG
    from solver_cpp_amd import gen
    out = {}
    vp = dict(outer_eps=1e-14, outer_maxloop=20000, outer_restart=21, inner_maxloop=11, inner_restart=12,
              inner_eps=0.0, inner_precision=32)
    gm = dict(outer_eps=1e-14, outer_maxloop=100000, outer_restart=21)
    cases = (("aniso2d_1024", "aniso2d", dict(spd=True, solver="vpcg", inner="cg",
                                               params=dict(outer_eps=1e-14, outer_maxloop=20000, inner_maxloop=10,
                                                           inner_precision=32),
                                               adjoint_solver="cg",
                                               adjoint_params=dict(outer_eps=1e-14, outer_maxloop=100000))),
             ("convdiff3d_pe08_192", "convdiff3d_pe08", dict(spd=False, solver="vpgmres", inner="gmres", params=vp,
                                                             adjoint_solver="gmres", adjoint_params=gm)),
             ("zipf_4m_rowscaled", "zipf", dict(spd=False, solver="vpgmres", inner="gmres", params=vp,
                                                adjoint_solver="gmres", adjoint_params=gm, row_scale=True)))
    for name, cfg, kw in cases:
        P = zipf_problem() if cfg == "zipf" else gen.problem(cfg)
        sysm = AG.SparseSystem(ctx, P["ptr"], P["col"], fp32=True, **kw)
        vals = torch.from_numpy(P["val"]).cuda().requires_grad_(True)
        b = torch.from_numpy(P["b"]).cuda().requires_grad_(True)
        xt = torch.from_numpy(P["x_target"]).cuda()
        # one untimed forward + backward first: the timed pair then measures a
        # solve with new values, not the first allocations of the row-scaled
        # copies and gradients
        (0.5 * ((sysm.solve(vals, b) - xt) ** 2).sum()).backward()
        vals.grad = None
        b.grad = None
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = sysm.solve(vals, b)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        (0.5 * ((x - xt) ** 2).sum()).backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        f, bw = sysm.last_forward, sysm.last_backward
        out[name] = {"forward_s": round(t1 - t0, 4), "grad_s": round(t2 - t1, 4), "forward_iters": f["iters"],
                     "forward_relres": f["relres_true"], "adjoint_iters": bw["iters"],
                     "adjoint_relres": bw["relres_true"], "n": len(P["ptr"]) - 1, "nnz": len(P["col"])}
        sysm.A.close()
        del vals, b, x
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(P, r0, a):
    """The oracle (this repository's FP64 C++/OpenMP restatement of the
    reference's loops) running the reference's own algorithm for the headline
    workload - VPGMRES(20) with an FP64 inner GMRES (vpgmres.cpp:840,
    innerMethods.cpp:74-111) - on the host cores: median of --cpu-runs whole
    solves, or with --cpu-steps a sample of that many outer steps per run,
    scaled to the full solve's outer step count."""
    import numpy as np
    from oracle import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (len(os.sched_getaffinity(0)) or 1)
    kw = dict(outer_restart=21, outer_maxloop=5000, outer_eps=1e-10, inner_maxloop=11, inner_restart=12,
              inner_eps=0.0, inner_precision=64, omp_threads=threads)
    if a.cpu_steps:
        kw["outer_maxloop"] = a.cpu_steps + 1  # count+1 >= maxloop stops after cpu_steps steps
    times, res = [], None
    for _ in range(max(1, a.cpu_runs)):
        t = time.perf_counter()
        xc, res, _ = O.solve(P["ptr"], P["col"], P["val"], P["b"], "vpgmres", "gmres", history=False, **kw)
        times.append(time.perf_counter() - t)
    med = float(np.median(times))
    if a.cpu_steps:
        # per outer step, scaled to the step count of the full solve (the GPU
        # FP64-inner solve's, which the oracle matches step for step)
        value = med / max(res["steps"], 1) * r0["steps"]
        sample = (f"oracle VPGMRES(20), FP64 inner GMRES (the reference's algorithm): {res['steps']} outer steps "
                  f"per run, median of {len(times)} runs {med:.2f} s, scaled to {r0['steps']} outer steps")
    else:
        value = med
        sample = (f"oracle VPGMRES(20), FP64 inner GMRES (the reference's algorithm), the whole solve: "
                  f"{res['steps']} outer steps, relres {res['relres_true']:.2e}, median of {len(times)} runs "
                  f"({', '.join(f'{t:.2f}' for t in times)} s)")
    # the host SpMV of the same matrix (reference blas::MtxVec_mult order)
    xs = np.random.default_rng(7).standard_normal(len(P["ptr"]) - 1)
    O.spmv(P["ptr"], P["col"], P["val"], xs)
    t = time.perf_counter()
    for _ in range(5):
        O.spmv(P["ptr"], P["col"], P["val"], xs)
    ts = (time.perf_counter() - t) / 5
    n, nnz = len(P["ptr"]) - 1, len(P["col"])
    return {"value": round(value, 3), "unit": "s", "cores": threads, "kind": "port", "sample": sample,
            "cpu_model": cpu_model(), "affinity_cpus": len(os.sched_getaffinity(0)),
            "spmv_fp64_GBps": round((12 * nnz + 20 * n + 4) / ts / 1e9, 1)}


def calib_copy(a):
    """One plain device-to-device copy of --calib-copy MiB (counter
    calibration: its FETCH_SIZE / WRITE_SIZE per byte moved scale the
    kernels' counters); returns the bytes it moved (read + written)."""
    import torch
    if not a.calib_copy:
        return 0
    src = torch.ones(a.calib_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    return 2 * src.numel() * 8


def hbm_ceiling(ctx, achieved, moved):
    """What plain streaming kernels reach on this box (1 GiB, best of four
    grids): the practical ceiling beside the published 8 TB/s, and the
    roofline kernel's rates as fractions of the read ceiling."""
    rd = ctx.stream_rate("read")
    out = {"what": "plain non-temporal 16-byte streaming kernels over 1 GiB (ks_bench_stream), best of 4 grids",
           "read_GBps": round(rd, 1), "copy_GBps": round(ctx.stream_rate("copy"), 1),
           "read_frac_of_peak": round(rd / HBM_PEAK_GBS, 4)}
    if achieved and moved:
        out["roofline_achieved_frac_of_read"] = round(achieved / rd, 4)
        out["roofline_format_frac_of_read"] = round(moved / rd, 4)
    return out


def spmv_gbps(A, n, nnz, prec, trans=False, reps=30):
    import torch
    dt = torch.float64 if prec == 64 else torch.float32
    xx = torch.randn(n, dtype=dt, device="cuda")
    yy = torch.empty_like(xx)
    A.spmv_bench(xx, yy, reps=3, trans=trans, precision=prec)
    ms = A.spmv_bench(xx, yy, reps=reps, trans=trans, precision=prec)
    sv = prec // 8
    ib = A.index_bytes(trans, prec)
    # the metric's byte model (int32 CSR), and the bytes of the format the
    # kernel reads (ib bytes per column index): its use of HBM bandwidth
    by32 = (sv + 4) * nnz + 4 * (n + 1) + 2 * sv * n
    by = by32 - (4 - ib) * nnz
    return {"us": round(ms * 1e3, 2), "GBps": round(by32 / (ms * 1e-3) / 1e9, 1),
            "frac_of_8TBps": round(by32 / (ms * 1e-3) / 8e12, 4), "index_bytes": ib,
            "format_GBps": round(by / (ms * 1e-3) / 1e9, 1), "format_frac": round(by / (ms * 1e-3) / 8e12, 4)}


def suite(ctx, a):
    """Per-configuration measurements: SpMV bandwidth (configs 2, 3, 5a, 5b),
    the configuration-1/2 solves, and the adjoint with its FD check
    (configurations 4 and 5)."""
    import numpy as np
    import torch
    from solver_cpp_amd import api, gen
    from solver_cpp_amd import autograd as AG

    def emit(d):
        print(json.dumps(d), flush=True)

    # --- SpMV bandwidth and parity solves ------------------------------
    for cfg in ("poisson2d", "convdiff3d", "poisson3d_sym", "poisson3d", "convdiff3d_pe08", "zipf"):
        P = gen.problem(cfg)
        n, nnz = len(P["ptr"]) - 1, len(P["col"])
        nonsym = cfg in ("convdiff3d", "convdiff3d_pe08", "zipf")
        A = api.Matrix(ctx, P["ptr"], P["col"], P["val"], fp32=True, transpose=nonsym)
        rec = {"suite": "spmv", "config": cfg, "n": n, "nnz": nnz,
               "fp64": spmv_gbps(A, n, nnz, 64), "fp32": spmv_gbps(A, n, nnz, 32)}
        if nonsym:
            rec["fp64_transpose"] = spmv_gbps(A, n, nnz, 64, trans=True)
        # the other SpMV layouts on the same matrix
        for lay in ("stream", "wave", "vector", "merge", "panels"):
            Al = api.Matrix(ctx, P["ptr"], P["col"], P["val"], fp32=True, layout=lay)
            rec[lay] = {"fp64": spmv_gbps(Al, n, nnz, 64), "fp32": spmv_gbps(Al, n, nnz, 32)}
            Al.close()
        emit(rec)
        b = torch.from_numpy(P["b"]).cuda()
        runs = []
        if cfg == "poisson2d":
            runs = [("cg", "NO", dict(outer_eps=1e-12, outer_maxloop=5000))]
        elif cfg == "convdiff3d":
            runs = [("bicg", "NO", dict(outer_eps=1e-10, outer_maxloop=20000)),
                    ("gmres", "NO", dict(outer_eps=1e-10, outer_maxloop=20000, outer_restart=31))]
        elif cfg == "poisson3d_sym":
            runs = [("kskipcg", "NO", dict(outer_eps=1e-10, outer_maxloop=20000, outer_kskip=4)),
                    ("cg", "NO", dict(outer_eps=1e-10, outer_maxloop=20000))]
        for outer, inner, kw in runs:
            api.solve(ctx, A, b, outer, inner, history=False, **kw)  # warm-up
            x, r, _ = api.solve(ctx, A, b, outer, inner, history=False, **kw)
            rec = {"suite": "solve", "config": cfg, "solver": outer, "status": r["status"], "iters": r["iters"],
                   "seconds": round(r["seconds"], 5), "relres_true": r["relres_true"],
                   "GBps": round(r["bytes"] / r["seconds"] / 1e9, 1)}
            if P.get("x_true") is not None:
                xt = torch.from_numpy(P["x_true"]).cuda()
                rec["err_vs_x_true"] = float(torch.linalg.norm(x - xt) / torch.linalg.norm(xt))
            emit(rec)
        A.close()

    # --- adjoint: configuration 4 (SPD) and 5 (nonsymmetric) -------------
    rng_dirs = {"aniso2d": 42, "convdiff3d_pe08": 54}
    for cfg in ("aniso2d", "convdiff3d_pe08"):
        P = gen.problem(cfg)
        n, nnz = len(P["ptr"]) - 1, len(P["col"])
        spd = cfg == "aniso2d"
        b = torch.from_numpy(P["b"]).cuda()
        xt = torch.from_numpy(P["x_target"]).cuda()
        if spd:
            # forward and adjoint to 1e-14: the FD comparison at 1e-6 needs
            # gradients accurate well below 1e-6 relative
            # forward VPCG (FP32 inner CG), backward CG + SDDMM; FD solves by CG
            fwd = dict(outer_eps=1e-14, outer_maxloop=20000, inner_maxloop=10, inner_precision=32)
            cg = dict(outer_eps=1e-14, outer_maxloop=100000)
            sysm = AG.SparseSystem(ctx, P["ptr"], P["col"], spd=True, solver="vpcg", inner="cg", fp32=True,
                                   params=fwd, adjoint_solver="cg", adjoint_params=dict(cg))
            fd_solver, fd_inner, fd_params = "cg", "NO", dict(cg)
        else:
            # forward mixed VPGMRES, backward FP64 GMRES on the explicit A^T
            fwd = dict(outer_eps=1e-14, outer_maxloop=20000, outer_restart=21, inner_maxloop=11,
                       inner_restart=12, inner_precision=32)
            sysm = AG.SparseSystem(ctx, P["ptr"], P["col"], spd=False, solver="vpgmres", inner="gmres",
                                   fp32=True, params=fwd, adjoint_solver="gmres",
                                   adjoint_params=dict(outer_eps=1e-14, outer_maxloop=100000, outer_restart=21))
            fd_solver, fd_inner, fd_params = "vpgmres", "gmres", dict(fwd)
        vals = torch.from_numpy(P["val"]).cuda().requires_grad_(True)
        bb = b.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = sysm.solve(vals, bb)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L = 0.5 * ((x - xt) ** 2).sum()
        L.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        rec = {"suite": "adjoint", "config": cfg, "n": n, "nnz": nnz, "forward_s": round(t1 - t0, 4),
               "backward_s": round(t2 - t1, 4), "forward_iters": sysm.last_forward["iters"],
               "forward_relres": sysm.last_forward["relres_true"],
               "adjoint_iters": sysm.last_backward["iters"],
               "adjoint_relres": sysm.last_backward["relres_true"]}
        # central FD along seeded directions (symmetric edge perturbations for the SPD case)
        rng = np.random.default_rng(rng_dirs[cfg])
        errs = []
        gv = vals.grad
        fdx = AG.CentralFD(ctx, P["ptr"], P["col"], vals.detach(), b, xt, 1e-6, fd_solver, fd_inner,
                           params=fd_params, spd=spd)
        t3 = time.perf_counter()
        for _ in range(a.fd_dirs):
            if spd:
                dk = rng.standard_normal(len(P["kappa"]))
                dv = torch.from_numpy(gen.aniso2d_values(P["grid"], P["grid"], dk)).cuda()
            else:
                dv = torch.from_numpy(rng.standard_normal(nnz)).cuda()
            fd = fdx(dv)
            ad = float(gv @ dv)
            errs.append(abs(fd - ad) / abs(ad))
        fdx.close()
        rec["fd_seconds"] = round(time.perf_counter() - t3, 2)
        rec["fd_dirs"] = a.fd_dirs
        rec["fd_max_rel_err"] = max(errs) if errs else None
        rec["fd_median_rel_err"] = float(np.median(errs)) if errs else None
        emit(rec)
        sysm.A.close()


if __name__ == "__main__":
    main()
