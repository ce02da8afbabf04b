#!/usr/bin/env python3
"""Benchmark of the headline metric: MDOF/s solved to 1e-8 relative residual, mixed precision vs
FP64 MG-PCG, with p50/p99 latency.

Workload (one step = one solve, right side resident in HBM to solution resident): 3D steady heat
conduction, hex8 on a structured 128^3 unit cube (2,113,407 free unknowns), per-cell conductivity
exp(N(0, 1)) from numpy's default_rng(1234), T = 0 on x = 0 and T = 1 on x = 1, insulated
elsewhere; 4-level geometric multigrid with Chebyshev(3)-Jacobi smoothing and a dense FP64 LDL^T
coarse level. `value` is the mixed solve (FP32 V-cycle inside the FP64 PCG); the plain FP64 MG-PCG
is run and reported beside it, and so is the mixed solve with the solver's automatic smoother degree
(3 on the finest level, 1 below). Usage:

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--no-cpu]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "MDOF/s solved to 1e-8 rel. residual, mixed vs FP64 MG-PCG; p50/p99 latency"
SETUP_THREADS = 8  # host threads of the serving loop that build admitted solves (half the box's 16-core share)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Hip:
    """The few HIP runtime calls the benchmark needs for device-resident vectors."""

    def __init__(self):
        path = None
        for cand in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
            try:
                self.lib = C.CDLL(cand)
                path = cand
                break
            except OSError:
                continue
        if path is None:
            raise RuntimeError("libamdhip64.so not found")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]

    def malloc(self, nbytes):
        p = C.c_void_p()
        if self.lib.hipMalloc(C.byref(p), nbytes):
            raise RuntimeError("hipMalloc failed")
        return p.value

    def h2d(self, dst, arr):
        if self.lib.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1):
            raise RuntimeError("hipMemcpy H2D failed")

    def d2h(self, arr, src):
        if self.lib.hipMemcpy(arr.ctypes.data_as(C.c_void_p), C.c_void_p(src), arr.nbytes, 2):
            raise RuntimeError("hipMemcpy D2H failed")

    def sync(self):
        self.lib.hipDeviceSynchronize()

    def free(self, p):
        self.lib.hipFree(C.c_void_p(p))


def get_profile(F, solver):
    return solver.profile()


def set_profile(F, solver, on):
    solver.set_profile(on)


def cheb_desc(degree, vector):
    """The smoother of the solves: an explicit degree, or the solver's automatic choice (degree 2
    on every level for vector problems and the assembled boundary)."""
    if degree > 0:
        return f"Chebyshev({degree})-Jacobi"
    return "Chebyshev(2)-Jacobi" if vector else "Chebyshev-Jacobi of degree 3 on the finest level, 1 below"


def heat_box(F, n):
    rng = np.random.default_rng(1234)
    k = np.exp(rng.standard_normal(n ** 3))
    return F.Box("heat_hex8", [n, n, n], [1.0, 1.0, 1.0], k, fixed={"x0": 0.0, "x1": 1.0}), k


def elastic_box(F, n):
    """Linear elasticity, tet10 on an n x n/2 x n/2 box of unit cells split into 6 tetrahedra per
    cell, E = 1, nu = 0.3, clamped at x = 0, unit body force in -z."""
    cells = [n, n // 2, n // 2]
    return F.Box("elastic_tet10", cells, [float(c) for c in cells], 1.0, 0.3, fixed={"x0": True}, body=(0.0, 0.0, -1.0))


def timed_solves(F, hip, solver, d_b, d_x, x0, steps, warmup, profile=False):
    """W untimed then K timed solves from x = 0; returns (latencies s, stats of the last, profile)."""
    for _ in range(warmup):
        hip.h2d(d_x, x0)
        solver.solve_device(d_b, d_x)
    lat = []
    st = None
    prof_s = prof_n = 0
    outer_s = outer_n = 0
    set_profile(F, solver, profile)
    hip.sync()
    t_all = time.perf_counter()
    for i in range(steps):
        hip.h2d(d_x, x0)  # the start value; outside the per-solve clock
        hip.sync()
        t0 = time.perf_counter()
        st = solver.solve_device(d_b, d_x)
        hip.sync()
        lat.append(time.perf_counter() - t0)
        if profile:
            p = get_profile(F, solver)
            prof_s += p.smoother_seconds
            prof_n += p.smoother_launches
            outer_s += p.outer_seconds
            outer_n += p.outer_launches
            bytes_ = (p.smoother_bytes, p.outer_bytes)
        log(f"  solve {i + 1}/{steps}: {lat[-1] * 1e3:.2f} ms, {st.iterations} iterations, residual {st.residual:.2e}")
    total = time.perf_counter() - t_all
    set_profile(F, solver, False)
    prof = None
    if profile and prof_n:
        prof = dict(smoother_s=prof_s / prof_n, smoother_bytes=bytes_[0], smoother_n=prof_n,
                    outer_s=outer_s / max(outer_n, 1), outer_bytes=bytes_[1], outer_n=outer_n)
    return lat, st, prof, total


def serve_requests(F, n_solves, seed):
    """The serving stream: refinement pairs (h, h/2) of heat tet4 or elasticity tet10 boxes (50/50),
    the coarse member's size log-uniform so that both members lie in about 5k-500k DOF. Boxes of
    2m x m x m cells, m the integer whose size is nearest the drawn one (tet4 heat: m in 13..31,
    fine 2m; tet10: m in 5..10, fine 2m: 4,900 and 3,888 free DOF at m = 13 and m = 4 fall below
    the stream's 5k floor, so heat starts at m = 14, 6,075 free DOF, and tet10 at m = 5, 7,260).
    Materials seeded."""
    rng = np.random.default_rng(seed)
    heat_m = list(range(14, 32))
    elas_m = list(range(5, 11))

    def dofs(kind, m):
        if kind == "heat_tet4":
            return (2 * m + 1) * (m + 1) ** 2
        return 3 * (4 * m + 1) * (2 * m + 1) ** 2

    reqs = []
    for p in range(n_solves // 2):
        kind = "heat_tet4" if rng.random() < 0.5 else "elastic_tet10"
        ms = heat_m if kind == "heat_tet4" else elas_m
        target = np.exp(rng.uniform(np.log(5e3), np.log(5e5 / 8)))
        m = min(ms, key=lambda v: abs(np.log(dofs(kind, v) / target)))
        for mm in (m, 2 * m):
            cells = [2 * mm, mm, mm]
            nt = 6 * int(np.prod(cells))
            length = [2.0, 1.0, 1.0]
            if kind == "heat_tet4":
                box = F.Box(kind, cells, length, np.exp(0.5 * rng.standard_normal(nt)), fixed={"x0": 0.0, "x1": 1.0})
            else:
                box = F.Box(kind, cells, length, rng.uniform(1.0, 2.0, nt), np.full(nt, 0.3), fixed={"x0": True},
                            body=(0.0, 0.0, -1.0))
            reqs.append((p, kind, mm, box, int(F.Model(box).free_list().size)))
    return reqs


def run_stream(F, reqs, arrivals, budget, max_active, batched=True):
    """Submits reqs at their arrival times (seconds from the start; None = all at once) and waits
    for all; returns (makespan s, latencies s, iterations, phases, memory, batching counts)."""
    srv = F.Server(hbm_budget=budget, max_active=max_active, setup_threads=SETUP_THREADS, batched=batched)
    t0 = time.perf_counter()
    ids = []
    opts = dict(method="multigrid", precision="fp32", tol=1e-8)
    for k, (_, kind, m, box, n) in enumerate(reqs):
        if arrivals is not None:
            dt = arrivals[k] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        ids.append(srv.submit(box, **opts))
    lat, its, phases = [], [], []
    for rid, (_, kind, m, box, n) in zip(ids, reqs):
        _, st, info = srv.wait(rid, n, timeout=600)
        lat.append(info.finish_time - info.submit_time)
        its.append(st.iterations)
        phases.append((info.admit_time - info.submit_time, info.start_time - info.admit_time,
                       info.finish_time - info.start_time))
    makespan = time.perf_counter() - t0
    mem = srv.memory()
    bst = srv.batching()
    srv.close()
    return makespan, np.array(lat), its, np.array(phases), mem, bst


def serving(F, n_solves, seed):
    reqs = serve_requests(F, n_solves, seed)
    total_dof = sum(r[4] for r in reqs)
    budget, max_active = 64 << 30, 32
    log(f"serving: {len(reqs)} solves ({total_dof} DOF), capacity run (all submitted at once), unbatched then batched")
    solo_span, _, _, _, _, _ = run_stream(F, reqs, None, budget, max_active, batched=False)
    cap_span, cap_lat, _, _, _, cap_bst = run_stream(F, reqs, None, budget, max_active)
    capacity = len(reqs) / cap_span
    rate = 0.8 * capacity
    rng = np.random.default_rng(seed + 1)
    pair_t = np.cumsum(rng.exponential(2.0 / rate, len(reqs) // 2))  # pairs arrive together
    arrivals = [pair_t[r[0]] for r in reqs]
    log(f"  capacity {capacity:.1f} solves/s ({total_dof / cap_span / 1e6:.1f} MDOF/s); Poisson arrivals at "
        f"{rate:.1f} solves/s")
    span, lat, its, ph, mem, bst = run_stream(F, reqs, arrivals, budget, max_active)
    out = {
        "workload": (f"{len(reqs)} solves in refinement pairs (h, h/2), heat tet4 or elasticity tet10 (50/50), "
                     f"{min(r[4] for r in reqs)}-{max(r[4] for r in reqs)} DOF, Poisson pair arrivals at 80% of the "
                     f"measured capacity (seed {seed}), mixed precision, tol 1e-8, HBM budget {budget >> 30} GiB, "
                     f"{max_active} solves admitted at once, {SETUP_THREADS} setup threads"),
        "solves": len(reqs), "total_dof": total_dof,
        "capacity_solves_s": round(capacity, 2), "capacity_mdof_s": round(total_dof / cap_span / 1e6, 2),
        "arrival_rate_solves_s": round(rate, 2),
        "mdof_s": round(total_dof / span / 1e6, 2), "makespan_s": round(span, 3),
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 2), "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 2),
        "mean_ms": round(1e3 * float(np.mean(lat)), 2), "max_iterations": int(max(its)),
        "mean_queue_ms": round(1e3 * float(ph[:, 0].mean()), 2), "mean_setup_ms": round(1e3 * float(ph[:, 1].mean()), 2),
        "mean_solve_ms": round(1e3 * float(ph[:, 2].mean()), 2),
        "latency": "submit to solution on the host, including setup (model, hierarchy, coarse factor)",
        "batching": {
            "capacity_unbatched_solves_s": round(len(reqs) / solo_span, 2),
            "capacity_batched_solves_s": round(capacity, 2),
            "stream": bst, "capacity_run": cap_bst,
            "solves_per_batch": round(bst["iterations"] / max(bst["batches"], 1), 2),
            "launches_per_batch": round(bst["launches"] / max(bst["batches"], 1), 1),
            "launches_per_batch_unbatched": round(bst["solo_launches"] / max(bst["batches"], 1), 1),
            "note": "continuous batching: each PCG iteration due of every running solve goes out in one batch, each "
                    "launch running one kernel (the same level of each V-cycle, the same PCG step) over all members; "
                    "one captured hipGraph per set of members; unbatched: each solve replays its own iteration graph "
                    "on its own stream"},
        "device_memory": {"arena_bytes": mem[0], "peak_bytes": mem[1], "allocations_outside_arena": mem[2],
                          "note": "every solve's device memory comes from one arena reserved at fem_server_new: "
                                  "no hipMalloc or hipFree while solves run"},
    }
    log(f"  stream: {out['mdof_s']} MDOF/s, latency p50 {out['p50_ms']} ms, p99 {out['p99_ms']} ms")
    return out


def measured_traffic(n):
    """HBM bytes per launch of the dominant kernel from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    kept under profiles/ (measured at this workload only), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_heat_hex8_{n}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return {"bytes_per_launch": t["bytes_per_launch"], "source": os.path.relpath(path, ROOT)}


def cpu_baseline(n_cpu, reference_tol=None):
    """The oracle (the reference method restated in C++: assembled CSR, line Gauss-Seidel GMG with
    Galerkin coarse operators and a sparse Cholesky coarsest level, FP64 PCG), one thread, on a
    box of the same workload; solve time to 1e-8 only (setup reported apart). With reference_tol,
    the oracle's PCG then continues from that solution to reference_tol, and its solution is
    returned beside the baseline (the FP64 oracle solution of the headline)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    rng = np.random.default_rng(1234)
    k = np.exp(rng.standard_normal(n_cpu ** 3))
    t0 = time.perf_counter()
    pr = O.Problem.heat([n_cpu] * 3, [1.0] * 3, k, fixed={"x0": 0.0, "x1": 1.0})
    t1 = time.perf_counter()
    s = O.Solver(pr, "multigrid")
    t2 = time.perf_counter()
    x, it, res = s.solve(pr.rhs(), tol=1e-8)
    t3 = time.perf_counter()
    nfree = pr.sizes.n_free
    x_ref = None
    if reference_tol is not None:
        x_ref, it_ref, res_ref = s.solve(pr.rhs(), x0=x, tol=reference_tol)
        log(f"  oracle FP64 MG-PCG continued to {res_ref:.1e} ({it_ref} more iterations, "
            f"{time.perf_counter() - t3:.1f} s)")
    return x_ref, {
        "assembly_s": round(t1 - t0, 3),
        "value": nfree / (t3 - t2) / 1e6,
        "unit": "MDOF/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"heat hex8 {n_cpu}^3 of the same field law ({nfree} free unknowns), reference method "
                   f"(line Gauss-Seidel GMG-PCG, FP64), {it} iterations to {res:.1e}; solve {t3 - t2:.2f} s, "
                   f"assembly {t1 - t0:.2f} s and setup {t2 - t1:.2f} s excluded; host {os.uname().machine}, "
                   f"nproc {os.cpu_count()}"),
    }


def tet10_smoother(prof):
    """The fused tet10 Chebyshev step (tile kernel and boundary pass) of a kernel timing."""
    gbs = prof["smoother_bytes"] / prof["smoother_s"] / 1e9
    return {"kernel": "KT10Tile<float> + KT10Bound<float> (fused tet10 operator and Chebyshev update, FP32, run by k_run)",
            "bytes_per_launch": int(prof["smoother_bytes"]), "us_per_launch": round(prof["smoother_s"] * 1e6, 2),
            "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}


def streaming_point(F, n):
    """The headline's dominant kernel (the FP32 hex8 Chebyshev step) on a heat box of n^3 cells,
    whose level-0 arrays (about 28 B per node) are several times the 256 MB last-level cache: its
    rate streams from HBM."""
    box, _ = heat_box(F, n)
    log(f"hbm streaming point: heat hex8 {n}^3, the finest Chebyshev step timed alone")
    s = F.Solver(box, method="multigrid", precision="fp32", levels=5, cheb_degree=3)
    tk = s.time_kernels(50)
    sec = tk.smoother_seconds / tk.smoother_launches
    gbs = tk.smoother_bytes / sec / 1e9
    log(f"  {tk.smoother_bytes / 1e6:.0f} MB in {sec * 1e6:.1f} us: {gbs:.0f} GB/s")
    return {"workload": f"heat hex8 {n}^3 (the same field law), the finest FP32 Chebyshev step",
            "bytes_per_launch": int(tk.smoother_bytes), "us_per_launch": round(sec * 1e6, 2),
            "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
            "measured": "HIP events around 50 back-to-back launches (fem_solver_time_kernels)"}


def run_workload(F, hip, box, levels, args, label, degree, auto_leg=False):
    """Mixed and FP64 MG-PCG on one box with Chebyshev degree `degree` (0: the solver's automatic
    choice): timed solves, the kernel timings, the error against an FP64 solve to 1e-13; with
    auto_leg, also the mixed solve with the automatic degree. Returns the result dict."""
    log(f"setup: {label}, mixed and FP64 hierarchies ({levels or 'automatic'} levels)")
    t0 = time.perf_counter()
    opts = dict(method="multigrid", tol=1e-8, levels=levels, cheb_degree=degree,
                coarse_galerkin=args.coarse == "galerkin", max_iter=2000)
    mixed = F.Solver(box, precision="fp32", **opts)
    t1 = time.perf_counter()
    fp64 = F.Solver(box, precision="fp64", **opts)
    t2 = time.perf_counter()
    bf16 = F.Solver(box, precision="bf16", **opts)
    t3 = time.perf_counter()
    nfree = mixed.n
    log(f"  {nfree} free unknowns on {mixed.device_name()}; setup mixed {t1 - t0:.2f} s, fp64 {t2 - t1:.2f} s")
    lv = []
    for l in range(32):
        try:
            lv.append(mixed.level_info(l))
        except F.FemError:
            break
    log("  levels (unknowns, lmax):", lv)
    b = mixed.rhs()
    x0 = np.zeros(nfree)
    d_b = hip.malloc(b.nbytes)
    d_x = hip.malloc(b.nbytes)
    hip.h2d(d_b, b)
    log("mixed (FP32 V-cycle, FP64 PCG): timed solves (replayed HIP graphs)")
    lat_m, st_m, _, _ = timed_solves(F, hip, mixed, d_b, d_x, x0, args.steps, args.warmup, profile=False)
    x_mixed = np.zeros(nfree)
    hip.d2h(x_mixed, d_x)
    log("mixed: the finest-level Chebyshev step and the outer q = A p, 100 launches each between HIP events")
    tk = mixed.time_kernels(100)
    prof = dict(smoother_s=tk.smoother_seconds / tk.smoother_launches, smoother_bytes=tk.smoother_bytes,
                smoother_n=tk.smoother_launches, outer_s=tk.outer_seconds / tk.outer_launches,
                outer_bytes=tk.outer_bytes, outer_n=tk.outer_launches)
    log(f"  smoother {prof['smoother_s'] * 1e6:.2f} us, outer apply {prof['outer_s'] * 1e6:.2f} us per launch")
    log("mixed with BF16-stored coefficients and D^-1 (FP32 V-cycle arithmetic, FP64 PCG): timed solves")
    lat_h, st_h, _, _ = timed_solves(F, hip, bf16, d_b, d_x, x0, args.steps, args.warmup, profile=False)
    x_bf16 = np.zeros(nfree)
    hip.d2h(x_bf16, d_x)
    tkh = bf16.time_kernels(100)
    prof_h = dict(smoother_s=tkh.smoother_seconds / tkh.smoother_launches, smoother_bytes=tkh.smoother_bytes)
    log(f"  bf16 smoother {prof_h['smoother_s'] * 1e6:.2f} us per launch ({prof_h['smoother_bytes']} bytes)")
    auto = None
    if auto_leg:
        log("mixed with the automatic smoother degree (3 on the finest level, 1 below): timed solves")
        ta = time.perf_counter()
        mixed_a = F.Solver(box, precision="fp32", **dict(opts, cheb_degree=0))
        ta = time.perf_counter() - ta
        lat_a, st_a, _, _ = timed_solves(F, hip, mixed_a, d_b, d_x, x0, args.steps, args.warmup, profile=False)
        x_auto = np.zeros(nfree)
        hip.d2h(x_auto, d_x)
        del mixed_a
        auto = (lat_a, st_a, x_auto, ta)
    log("FP64 MG-PCG: timed solves")
    lat_d, st_d, _, _ = timed_solves(F, hip, fp64, d_b, d_x, x0, args.steps, args.warmup, profile=False)
    x_fp64 = np.zeros(nfree)
    hip.d2h(x_fp64, d_x)
    fp64.set_tol(1e-13)
    hip.h2d(d_x, x0)
    st_ref = fp64.solve_device(d_b, d_x)
    x_ref = np.zeros(nfree)
    hip.d2h(x_ref, d_x)
    hip.free(d_b)
    hip.free(d_x)
    err_mixed = float(np.linalg.norm(x_mixed - x_ref) / np.linalg.norm(x_ref))
    err_fp64 = float(np.linalg.norm(x_fp64 - x_ref) / np.linalg.norm(x_ref))
    err_bf16 = float(np.linalg.norm(x_bf16 - x_ref) / np.linalg.norm(x_ref))
    log(f"  error vs FP64 1e-13 solve: mixed {err_mixed:.2e}, bf16 {err_bf16:.2e}, fp64 {err_fp64:.2e}; reference took "
        f"{st_ref.iterations} it")

    def leg(lat, st, err, setup):
        mean = float(np.mean(lat))
        return {"mdof_s": round(nfree / mean / 1e6, 2), "ms_mean": round(1e3 * mean, 3),
                "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
                "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3), "iterations": st.iterations,
                "residual": st.residual, "true_residual": st.true_residual, "restarts": st.restarts,
                "error_vs_fp64": err, "setup_s": round(setup, 3),
                # the algorithmic bytes of every launch of one solve (the library's ledger) over the
                # mean solve time, against the 8 TB/s HBM peak
                "solve_bytes": int(st.hbm_bytes), "solve_gbs": round(st.hbm_bytes / mean / 1e9, 1),
                "solve_frac": round(st.hbm_bytes / mean / 1e9 / HBM_PEAK_GBS, 4)}

    res = {"free_dof": nfree, "levels": [v[0] for v in lv], "mixed": leg(lat_m, st_m, err_mixed, t1 - t0),
           "mixed_bf16": leg(lat_h, st_h, err_bf16, t3 - t2),
           "fp64": leg(lat_d, st_d, err_fp64, t2 - t1),
           "fp64_reference": f"FP64 MG-PCG to 1e-13 relative residual ({st_ref.iterations} iterations)"}
    if auto is not None:
        lat_a, st_a, x_auto, ta = auto
        res["mixed_automatic_degree"] = leg(lat_a, st_a, float(np.linalg.norm(x_auto - x_ref) / np.linalg.norm(x_ref)), ta)
        res["mixed_automatic_degree"]["smoother"] = cheb_desc(0, False)
    res["speedup_mixed_vs_fp64"] = round(res["mixed"]["mdof_s"] / res["fp64"]["mdof_s"], 3)
    res["speedup_bf16_vs_fp64"] = round(res["mixed_bf16"]["mdof_s"] / res["fp64"]["mdof_s"], 3)
    res["mixed_bf16"]["finest_smoother"] = {"bytes_per_launch": int(prof_h["smoother_bytes"]),
                                            "us_per_launch": round(prof_h["smoother_s"] * 1e6, 2),
                                            "achieved_gbs": round(prof_h["smoother_bytes"] / prof_h["smoother_s"] / 1e9, 1)}
    res["_x_ref"] = x_ref
    res["_x"] = {"mixed": x_mixed, "mixed_bf16": x_bf16, "fp64": x_fp64}
    if auto is not None:
        res["_x"]["mixed_automatic_degree"] = auto[2]
    return res, prof, lat_m


def device_assembly(F, box, k, solver, ptr, idx, val):
    """fem_assemble on the heat box's mesh in the model's numbering (the reference's first-seen
    points, cells x fastest with VTK corners): the pattern and free list equal the export's bit for
    bit, the values and the right side equal it to the export's probe rounding; the host and device
    times of the assembly."""
    n = list(box.n)
    h = [box.length[d] / n[d] for d in range(3)]
    m = F.Model(box)
    pg = m.point_grid().astype(np.int64)
    at = np.full((n[0] + 1) * (n[1] + 1) * (n[2] + 1), -1, np.int64)
    at[pg[:, 0] + (n[0] + 1) * (pg[:, 1] + (n[1] + 1) * pg[:, 2])] = np.arange(pg.shape[0])
    i, j, kk = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    i, j, kk = (v.transpose(2, 1, 0).ravel() for v in (i, j, kk))  # cells x fastest
    bits = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    conn = np.stack([at[(i + a) + (n[0] + 1) * ((j + b) + (n[1] + 1) * (kk + c))] for a, b, c in bits], axis=1)
    xyz = pg * np.array(h)
    fixed = (pg[:, 0] == 0) | (pg[:, 0] == n[0])
    value = np.where(pg[:, 0] == n[0], 1.0, 0.0)
    mesh = F.Mesh("heat_hex8", xyz, conn.astype(np.uint32), k, fixed.astype(np.uint8), value)
    t0 = time.perf_counter()
    a = F.Assembly(mesh)
    t1 = time.perf_counter()
    aptr, aidx, aval, arhs, afree = a.arrays()
    st = a.stats
    same_pattern = bool(np.array_equal(afree, solver.free_list()) and np.array_equal(aptr, ptr.astype(np.uint64))
                        and np.array_equal(aidx, idx))
    val_err = float(np.abs(aval - val).max() / np.abs(val).max())
    b = solver.rhs()
    rhs_err = float(np.abs(arhs - b).max() / np.abs(b).max())
    cells = conn.shape[0]
    # algorithmic bytes of the device part: the element matrices written and read back by the
    # gather (8 B a slot), the gather maps (4 B a slot, 8 B an entry), the values written, the
    # reduced copy and right side (their arrays once), the mesh read by the element kernel
    slots = 64 * cells
    nbytes = (slots * (8 + 8 + 4) + 8 * int(st.full_entries) * 2 + 8 * cells * 8 * 2 + 12 * int(st.entries)
              + 8 * cells * 4 + 24 * pg.shape[0])
    log(f"  device assembly: {int(st.entries)} reduced entries; host {st.host_seconds:.3f} s, device "
        f"{st.device_seconds * 1e3:.2f} ms ({nbytes / st.device_seconds / 1e9:.0f} GB/s algorithmic); pattern equal to "
        f"the export: {same_pattern}; values {val_err:.1e}, right side {rhs_err:.1e} from the export")
    out = {"workload": (f"fem_assemble of the same box's hex8 mesh ({cells} cells, {pg.shape[0]} points) in the "
                        "model's numbering: pattern, free list and gather maps on the host, element matrices, "
                        "gathers, reduced matrix and right side on the device"),
           "entries": int(st.entries), "host_s": round(st.host_seconds, 3),
           "device_ms": round(st.device_seconds * 1e3, 3), "call_s": round(t1 - t0, 3),
           "device_bytes": int(nbytes), "device_gbs": round(nbytes / st.device_seconds / 1e9, 1),
           "device_frac": round(nbytes / st.device_seconds / 1e9 / HBM_PEAK_GBS, 4),
           "pattern_equal_to_export": same_pattern, "values_rel_diff_vs_export": val_err,
           "rhs_rel_diff_vs_export": rhs_err}
    a.close()
    return out


def assembled_leg(F, hip, solver, box, x_ref, args, degree, k=None):
    """The assembled boundary on the same heat box: the model's reduced CSR (fem_solver_matrix, as
    the reference's caller assembles it), its grid unknowns, and fem_solver_new_csr with the
    reference's grid hierarchy (coarsen and Galerkin products down to 3000 unknowns), mixed
    precision, solved on the device like the box solver."""
    log("assembled boundary: the model's reduced CSR and grid unknowns into fem_solver_new_csr")
    t0 = time.perf_counter()
    ptr, idx, val = solver.matrix()
    n = list(box.n)
    planes = [(box.length[d] / n[d]) * np.arange(n[d] + 1, dtype=np.float64) for d in range(3)]
    pts = F.Model(box).point_grid()[solver.free_list()]
    t1 = time.perf_counter()
    csr = F.CsrSolver(ptr, idx, val, "grid", F.Grid(planes, pts, np.zeros(solver.n, np.uint8)), method="multigrid",
                      precision="fp32", tol=1e-8, max_iter=2000, cheb_degree=degree)
    t2 = time.perf_counter()
    lv = []
    for l in range(32):
        try:
            lv.append(csr.level_info(l)[0])
        except F.FemError:
            break
    log(f"  {int(ptr[-1])} entries; export {t1 - t0:.2f} s, setup {t2 - t1:.2f} s; levels {lv}")
    assembly = device_assembly(F, box, k, solver, ptr, idx, val) if k is not None else None
    b = solver.rhs()
    d_b = hip.malloc(b.nbytes)
    d_x = hip.malloc(b.nbytes)
    hip.h2d(d_b, b)
    x0 = np.zeros(solver.n)
    lat, st, _, _ = timed_solves(F, hip, csr, d_b, d_x, x0, args.steps, args.warmup, profile=False)
    x = np.zeros(solver.n)
    hip.d2h(x, d_x)
    hip.free(d_b)
    hip.free(d_x)
    err = float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref))
    return {"mdof_s": round(solver.n / float(np.mean(lat)) / 1e6, 2), "ms_mean": round(1e3 * float(np.mean(lat)), 3),
            "iterations": st.iterations, "residual": st.residual, "true_residual": st.true_residual,
            "error_vs_fp64": err, "entries": int(ptr[-1]), "levels": lv, "export_s": round(t1 - t0, 3),
            "setup_s": round(t2 - t1, 3), "device_assembly": assembly,
            "workload": "the same heat box through fem_solver_new_csr: CSR SpMV operators on every level (FP32 "
                        "V-cycle, FP64 PCG), the reference's coarsen and Galerkin levels down to 3000 unknowns, "
                        f"{cheb_desc(degree, True)}, dense LDLT coarse, tol 1e-8"}


FP64_MFMA_PEAK_TFLOPS = 78.6  # MI355X dense FP64 matrix-core peak


def many_leg(F, box, k, seed=11):
    """k right sides against one setup: k solves one after another (fem_solver_solve), then the same
    k at once (fem_solver_solve_many: shadow solvers sharing the hierarchy, their iterations
    batched). Host vectors on both sides; wall time of the calls."""
    s = F.Solver(box, method="multigrid", precision="fp32", tol=1e-8)
    rng = np.random.default_rng(seed)
    b = s.rhs()
    scale = np.linalg.norm(b) / np.sqrt(s.n)
    B = np.array([(1.0 + 0.1 * i) * b + 0.05 * scale * rng.standard_normal(s.n) for i in range(k)])
    s.solve_many(B)  # the shadows are made on the first call; warm both paths
    s.solve(B[0])
    t0 = time.perf_counter()
    seq = [s.solve(B[i]) for i in range(k)]
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    X, stats = s.solve_many(B)
    t_many = time.perf_counter() - t0
    equal = all(np.array_equal(X[i], seq[i][0]) for i in range(k))
    return {"right_sides": k, "free_dof": int(s.n), "iterations": [int(st.iterations) for st in stats],
            "one_after_another_ms": round(1e3 * t_seq, 2), "together_ms": round(1e3 * t_many, 2),
            "one_after_another_mdof_s": round(k * s.n / t_seq / 1e6, 2), "together_mdof_s": round(k * s.n / t_many / 1e6, 2),
            "speedup": round(t_seq / t_many, 3), "bitwise_equal_to_alone": bool(equal),
            "note": "wall time of the calls with host vectors (PCIe copies on both sides)"}


def dense_factor_leg(F, hip, n, seed=5):
    """The dense FP64 LDL^T alone (the coarse level's and the direct method's factor, through
    fem_dense_solve_device): a random SPD matrix of n unknowns, the call timed from the host
    (factor, explicit inverse and a refined solve; the factor and inverse are n^3 flops, the
    solves n^2). Returns the timing and the achieved FP64 TFLOP/s against the matrix-core peak."""
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((n, 16))
    a = np.diag(2.0 + rng.uniform(0.0, 1.0, n))
    a += 0.1 * (m @ m.T) / 16
    b = rng.standard_normal(n)
    da, db, dx = F.DeviceArray(a), F.DeviceArray(b), F.DeviceArray(np.zeros(n))
    err = C.create_string_buffer(512)
    times = []
    for rep_ in range(3):
        hip.sync()
        t0 = time.perf_counter()
        if F.lib().fem_dense_solve_device(n, da.ptr, db.ptr, dx.ptr, err, 512):
            raise RuntimeError(err.value.decode())
        hip.sync()
        times.append(time.perf_counter() - t0)
    x = dx.get()
    resid = float(np.linalg.norm(b - a @ x) / np.linalg.norm(b))
    np_ = (n + 63) // 64 * 64
    flops = 3 * np_ ** 3 / 3.0  # the factor (n^3 / 3) and the two products of the inverse (n^3 / 3 each)
    t = min(times)
    log(f"dense LDL^T n={n}: {t * 1e3:.1f} ms, {flops / t / 1e12:.2f} TFLOP/s, residual {resid:.1e}")
    return {"n": n, "seconds": round(t, 5), "flops": flops, "tflops": round(flops / t / 1e12, 3),
            "frac_of_fp64_mfma_peak": round(flops / t / 1e12 / FP64_MFMA_PEAK_TFLOPS, 4), "residual": resid,
            "note": "host time of fem_dense_solve_device (factor, explicit inverse, refined solve), best of 3"}


def multifrontal_leg(F, hip, n):
    """The direct method beyond the dense factor: the multifrontal LDL^T (fem_solver_new_csr with
    the direct method) of the heat hex8 box of n^3 cells (the field law of the headline), its matrix
    exported by fem_solver_matrix. Reports the setup (nested dissection, symbolic analysis, the
    fronts' factor on the device), the solve, the true residual, and the factor's counts (the
    library's host symbolic analysis)."""
    box, _ = heat_box(F, n)
    sb = F.Solver(box, method="multigrid")
    ptr, idx, val = sb.matrix()
    pg = F.Model(box).point_grid()[sb.free_list()]
    planes = [np.arange(n + 1, dtype=np.float64) / n for _ in range(3)]
    g = F.Grid(planes, pg.astype(np.uint32), np.zeros(len(pg), np.uint8))
    b = sb.rhs()
    del sb
    xyz = np.stack([planes[d][pg[:, d]] for d in range(3)], axis=1)
    t0 = time.perf_counter()
    _, entries, flops, largest, supernodes = F.symbolic(ptr, idx, xyz)
    t_sym = time.perf_counter() - t0
    hip.sync()
    t0 = time.perf_counter()
    s = F.CsrSolver(ptr, idx, val, "grid", g, method="direct", direct_solver="multifrontal")
    hip.sync()
    t_setup = time.perf_counter() - t0
    sts = [s.solve(b)[1] for _ in range(3)]
    st = min(sts, key=lambda q: q.seconds)
    log(f"multifrontal heat {n}^3 ({len(b)} unknowns): setup {t_setup:.3f} s (host symbolic {t_sym:.3f} s), "
        f"solve {st.seconds * 1e3:.2f} ms, residual {st.true_residual:.1e}")
    return {"workload": f"heat hex8 {n}^3 box matrix (exported by fem_solver_matrix), direct method, multifrontal LDL^T",
            "unknowns": int(len(b)), "factor_entries": int(entries), "factor_flops": float(flops),
            "supernodes": int(supernodes), "largest_front_entries": int(largest),
            "setup_s": round(t_setup, 4), "host_symbolic_s": round(t_sym, 4),
            "solve_ms": round(st.seconds * 1e3, 3), "true_residual": float(st.true_residual),
            "factor_tflops_incl_host": round(flops / t_setup / 1e12, 4),
            "note": "setup is the host time of fem_solver_new_csr (order, symbolic analysis, level lists, the "
                    "device factor); solve is the device time of one solve (best of 3)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--cpu-n", type=int, default=128,
                    help="cells per axis of the oracle's heat box (at --n it also gives the error against the "
                         "oracle's FP64 solution to 1e-13)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--degree", type=int, default=None,
                    help="Chebyshev degree on every level (0: automatic); default: 3 for the heat box as its "
                         "configuration names, automatic for the elasticity boxes")
    ap.add_argument("--coarse", choices=["galerkin", "mean"], default="galerkin")
    ap.add_argument("--elastic-n", type=int, default=60,
                    help="cells along x of the tet10 elasticity box reported beside the headline (0: skip)")
    ap.add_argument("--target-n", type=int, default=70,
                    help="cells along x of the 2M-DOF tet10 target box (70: 2,117,220 unknowns; 0: skip)")
    ap.add_argument("--stream-n", type=int, default=256,
                    help="cells per axis of the heat box whose finest smoother step is timed past the last-level "
                         "cache (0: skip)")
    ap.add_argument("--inclusions-n", type=int, default=80,
                    help="cells along x of the nearly incompressible inclusion box (0: skip)")
    ap.add_argument("--inclusions", type=int, default=200, help="spheres in the inclusion box")
    ap.add_argument("--serve-n", type=int, default=256, help="solves of the serving stream (0: skip)")
    ap.add_argument("--many", type=int, default=8, help="right sides of the multi-RHS leg on the tet10 block (0: skip)")
    ap.add_argument("--no-assembled", action="store_true", help="skip the assembled (CSR) boundary leg of the heat box")
    ap.add_argument("--dense", type=str, default="4335,20000",
                    help="sizes of the dense LDL^T factor timed alone (comma-separated; empty: skip)")
    ap.add_argument("--dense-only", action="store_true", help="time the dense factors of --dense and stop")
    ap.add_argument("--multifrontal", type=str, default="34,50",
                    help="cells per axis of the heat boxes whose direct solve by the multifrontal factor is timed "
                         "(comma-separated; empty: skip)")
    ap.add_argument("--calibrate", action="store_true",
                    help="run one plain 256 MiB device copy first (the yardstick of HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (every box here has one GPU; "
              "independent solves would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    import femahoy_amd as F

    hip = Hip()
    if args.calibrate:
        cal = 256 << 20
        d_c0, d_c1 = hip.malloc(cal), hip.malloc(cal)
        err = C.create_string_buffer(256)
        F.lib().fem_test_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_char_p, C.c_size_t]
        if F.lib().fem_test_copy(C.c_void_p(d_c0), C.c_void_p(d_c1), cal, err, 256):
            raise RuntimeError(err.value.decode())
        log(f"calibration: one copy kernel of {cal} bytes read + {cal} bytes written")
        hip.free(d_c0)
        hip.free(d_c1)

    dense = [dense_factor_leg(F, hip, int(v)) for v in args.dense.split(",") if v.strip()]
    if args.dense_only:
        print(json.dumps({"dense_factor": dense}), flush=True)
        return
    multifrontal = [multifrontal_leg(F, hip, int(v)) for v in args.multifrontal.split(",") if v.strip()]

    # the heat configuration names Chebyshev(3)-Jacobi; the elasticity ones name no degree
    hdeg = 3 if args.degree is None else args.degree
    vdeg = 0 if args.degree is None else args.degree
    n = args.n
    box, heat_k = heat_box(F, n)
    heat, prof, lat_m = run_workload(F, hip, box, args.levels, args, f"heat hex8 {n}^3", hdeg,
                                     auto_leg=args.degree is None)
    value = heat["mixed"]["mdof_s"]
    x_ref = heat.pop("_x_ref")
    heat_x = heat.pop("_x")
    if not args.no_assembled:
        heat["assembled_csr"] = assembled_leg(F, hip, F.Solver(box, method="multigrid", levels=args.levels), box,
                                              x_ref, args, vdeg, k=heat_k)

    roofline = None
    traffic = measured_traffic(n)
    if prof:
        achieved = prof["smoother_bytes"] / prof["smoother_s"] / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic["bytes_per_launch"] if traffic else None,
                    "kernel": "k_march_hex8<float, MARCH_CHEB> (finest-level Chebyshev step: hex8 operator fused "
                              "with the update, FP32)",
                    "bytes_per_launch": int(prof["smoother_bytes"]),
                    "us_per_launch": round(prof["smoother_s"] * 1e6, 2), "launches": int(prof["smoother_n"]),
                    "measured": "HIP events on the solver stream around 100 back-to-back launches of the kernel "
                                "on the solver's own level-0 vectors (fem_solver_time_kernels)",
                    "working_set": ("cache-resident: the launch's 60 MB of level-0 arrays stay in the 256 MB "
                                    "last-level cache across back-to-back launches, so this is not an HBM-streaming "
                                    "rate; see hbm_streaming for the same kernel on a working set past the cache, "
                                    "and solve_frac for the whole solve"),
                    "solve_frac": heat["mixed"]["solve_frac"], "solve_bytes": heat["mixed"]["solve_bytes"],
                    "solve_gbs": heat["mixed"]["solve_gbs"],
                    "traffic_source": traffic["source"] if traffic else None,
                    "outer_apply": {"kernel": "k_march_hex8<double, MARCH_PCG_Q> (FP64 operator, q = A p, p.q partials)",
                                    "bytes_per_launch": int(prof["outer_bytes"]),
                                    "us_per_launch": round(prof["outer_s"] * 1e6, 2),
                                    "achieved": round(prof["outer_bytes"] / prof["outer_s"] / 1e9, 1)}}

    if roofline is not None and args.stream_n > 0:
        roofline["hbm_streaming"] = streaming_point(F, args.stream_n)

    elastic = None
    if args.elastic_n > 0:
        en = args.elastic_n
        ebox = elastic_box(F, en)
        elastic, eprof, _ = run_workload(F, hip, ebox, 0, args, f"elasticity tet10 {en}x{en // 2}x{en // 2}", vdeg)
        elastic.pop("_x_ref")
        elastic.pop("_x")
        elastic["workload"] = (f"linear elasticity, tet10 on {en}x{en // 2}x{en // 2} unit cells (6 Kuhn tets per cell), "
                               "E=1, nu=0.3, clamped at x=0, body force (0,0,-1); p-MG tet10->tet4 then geometric tet4 "
                               f"levels, dense FP64 LDLT coarse, {cheb_desc(vdeg, True)}, tol 1e-8")
        if eprof:
            elastic["finest_smoother"] = tet10_smoother(eprof)
        if args.many > 0:
            log(f"many right sides: {args.many} against one tet10 setup")
            elastic["many_right_sides"] = many_leg(F, ebox, args.many)

    target = None
    if args.target_n > 0:
        tn = args.target_n
        tbox = elastic_box(F, tn)
        target, tprof, _ = run_workload(F, hip, tbox, 0, args, f"elasticity tet10 {tn}x{tn // 2}x{tn // 2}", vdeg)
        target.pop("_x_ref")
        target.pop("_x")
        target["workload"] = (f"the 2M-DOF tet10 target: linear elasticity, tet10 on {tn}x{tn // 2}x{tn // 2} unit cells, "
                              "E=1, nu=0.3, clamped at x=0, body force (0,0,-1); p-MG tet10->tet4 then geometric tet4 "
                              f"levels, dense FP64 LDLT coarse, {cheb_desc(vdeg, True)}, tol 1e-8")
        if tprof:
            target["finest_smoother"] = tet10_smoother(tprof)
        frac = target["finest_smoother"]["frac"] if tprof else None
        target["targets"] = {
            "speedup_mixed_vs_fp64": {"want": 2.0, "got": target["speedup_mixed_vs_fp64"],
                                      "met": target["speedup_mixed_vs_fp64"] >= 2.0},
            "fused_step_roofline_frac": {"want": 0.6, "got": frac, "met": frac is not None and frac >= 0.6},
            "solve_roofline_frac": {"want": 0.6, "got": target["mixed"]["solve_frac"],
                                    "met": target["mixed"]["solve_frac"] >= 0.6},
            "error_vs_fp64": {"want": 1e-8, "got": target["mixed"]["error_vs_fp64"],
                              "met": target["mixed"]["error_vs_fp64"] <= 1e-8},
        }

    inclusions = None
    if args.inclusions_n > 0:
        ni = args.inclusions_n
        ibox, inside = F.inclusion_box(ni, args.inclusions, 7)
        inclusions, _, _ = run_workload(F, hip, ibox, 0, args, f"inclusions tet10 {ni}x{ni // 2}x{ni // 2}", vdeg)
        inclusions.pop("_x_ref")
        inclusions.pop("_x")
        inclusions["workload"] = (f"nearly incompressible heterogeneous elasticity, tet10 on {ni}x{ni // 2}x{ni // 2} unit "
                                  f"cells, nu=0.49 matrix (E=1) with {args.inclusions} seeded spheres (radius 0.02-0.05 "
                                  "of the box length, E=1e3, nu=0.3; a tet is inside when its centroid is), clamped at "
                                  "x=0, traction (1,0,0) on x=1; p-MG tet10->tet4 then geometric tet4 levels, dense FP64 "
                                  f"LDLT coarse, {cheb_desc(vdeg, True)}, tol 1e-8 with refinement against the FP64 true "
                                  "residual")
        inclusions["stiff_fraction"] = round(float(inside.mean()), 4)

    serve = serving(F, args.serve_n, 4321) if args.serve_n > 0 else None

    cpu = None
    if not args.no_cpu:
        log(f"cpu baseline: oracle on heat hex8 {args.cpu_n}^3, 1 thread")
        x_oracle, cpu = cpu_baseline(args.cpu_n, reference_tol=1e-13 if args.cpu_n == n else None)
        log("  " + cpu["sample"])
        if x_oracle is not None:
            # every heat leg against the oracle's FP64 MG-PCG solution to 1e-13 (the same vector
            # layout: the free unknowns in the reference order)
            on = float(np.linalg.norm(x_oracle))
            for name, xv in heat_x.items():
                heat[name]["error_vs_oracle"] = float(np.linalg.norm(xv - x_oracle) / on)
            heat["oracle_reference"] = ("the oracle's FP64 MG-PCG (line Gauss-Seidel GMG, Galerkin levels, sparse "
                                        "Cholesky coarsest level) on the same box to 1e-13 relative residual")
            heat["error_fp64_reference_vs_oracle"] = float(np.linalg.norm(x_ref - x_oracle) / on)
        da = heat.get("assembled_csr", {}).get("device_assembly")
        if da and args.cpu_n == n:
            # the oracle's model of the same box: numbering, pattern, add_element, reduced right side
            da["oracle_cpu_model_s"] = cpu["assembly_s"]
            log("  error vs the oracle: " + ", ".join(f"{k} {heat[k]['error_vs_oracle']:.2e}" for k in heat_x))

    config = {"workload": (f"heat hex8 {n}^3 unit cube, k = exp(N(0,1)) per cell (seed 1234), T=0 at x=0 and T=1 at "
                           f"x=1, {args.levels}-level GMG ({args.coarse} coarse operators), {cheb_desc(hdeg, False)}, "
                           "dense FP64 LDLT coarse, tol 1e-8")}
    config.update(heat)
    if elastic:
        config["elasticity_tet10"] = elastic
    if target:
        config["elasticity_tet10_2m"] = target
    if inclusions:
        config["inclusions_tet10"] = inclusions
    if serve:
        config["serving"] = serve
    if dense:
        config["dense_factor"] = dense
    if multifrontal:
        config["direct_multifrontal"] = multifrontal
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "MDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": heat["mixed"]["ms_mean"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 PCG + f32 V-cycle",
        "data": "synthetic (seeded lognormal conductivity; uniform elastic material)",
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
