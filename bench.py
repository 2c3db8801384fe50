#!/usr/bin/env python
"""Benchmark: material-point updates/s of the spectral-basic FCC phenopowerlaw
tension run on one MI355X.

Workload: 64^3 cells, 200-grain seeded periodic Voronoi, uniform random
orientations, FCC phenopowerlaw + Hooke, isostrain, uniaxial tension
dot_F_xx = 1e-3/s with all other P components zero, 100 increments to 5 %
(t = 50 s), FP64.  A step is one increment (forward, NGMRES solution,
updateCoords).  W warm-up increments run first; the next K increments are
timed.  One material-point update is one per-cell constitutive evaluation;
updates/s = cells x residual evaluations / wall time of the timed increments.

The CPU baseline is the oracle (a C/OpenMP restatement of the same algorithm)
on a bounded sample of the same workload, timed on this host's cores.
"""
import argparse
import copy
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0        # MI355X HBM3E
# HBM bytes per stress-kernel launch on this workload: rocprofv3 --pmc
# FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (factors from a plain device copy), see
# profiles/kernels_64cube_fcc_tension.md
STRESS_TRAFFIC_BYTES = 1882666106.9


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=98, help="timed increments")
    ap.add_argument("--warmup", type=int, default=2, help="untimed increments before the timed ones")
    ap.add_argument("--cells", type=int, default=64)
    ap.add_argument("--grains", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--increments", type=int, default=100, help="increments of the full load case")
    ap.add_argument("--strain", type=float, default=0.05)
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-cells", type=int, default=64, help="grid of the CPU baseline (the headline grid)")
    ap.add_argument("--cpu-increments", type=int, default=2,
                    help="timed increments of the CPU baseline, after the same warm-up increments as the GPU")
    ap.add_argument("--traffic-bytes", type=float, default=STRESS_TRAFFIC_BYTES,
                    help="measured HBM bytes per stress-kernel launch (rocprofv3 --pmc); 0 for none")
    ap.add_argument("--adjoint", type=int, default=1,
                    help="also measure adjoint / forward cost on the 32^3 differentiation workload")
    ap.add_argument("--adjoint-rtol", type=float, default=1e-6,
                    help="adjoint GMRES tolerance (1e-6: gradient within 2e-7 of the 1e-12 solve at 32^3)")
    ap.add_argument("--batch", type=int, default=256,
                    help="RVEs of the batched forward + adjoint measurement (0: skip)")
    ap.add_argument("--batch-threads", type=int, default=16)
    ap.add_argument("--lockstep-threads", type=int, default=256,
                    help="RVEs in flight of the lockstep batched forward (one host thread each; 0: skip)")
    ap.add_argument("--only-batch", type=int, default=0,
                    help="only run the batched measurement and print it as one JSON line")
    ap.add_argument("--only-adjoint", type=int, default=0,
                    help="only run the adjoint / forward measurement and print it as one JSON line")
    ap.add_argument("--mixed-batch", type=int, default=64,
                    help="RVEs of the 48^3 BCC mixed-load batched forward + checkpointed adjoint (0: skip)")
    ap.add_argument("--mixed-increments", type=int, default=50)
    ap.add_argument("--only-mixed", type=int, default=0, help="only the 48^3 BCC mixed-load batch, as one JSON line")
    ap.add_argument("--bcc", type=int, default=1,
                    help="also time the same workload with the 24-system BCC phase")
    ap.add_argument("--schemes", type=int, default=1,
                    help="also time the spectral_polarization and spectral_Galerkin schemes on the same workload")
    ap.add_argument("--only-schemes", type=int, default=0, help="only the other schemes' timings")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a plain device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def copy_calibration(mib):
    """A plain copy kernel of known size: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
    it gives the factor between counter readings and bytes moved."""
    import torch
    n = mib * 1024 * 1024 // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        y = x.clone()
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8,
                      "checksum": float(y[-1])}))


def workload(cells, grains, seed, increments, strain, phase=None):
    from damask_amd import config, workloads
    grid, mat = workloads.polycrystal([cells] * 3, grains, seed, phase=phase)
    t = strain / 1.0e-3
    _, lcs = config.parse_load(workloads.tension_x(t, increments))
    return grid, config.parse_material(mat), lcs[0]


def sub_loadcase(lc, n):
    """first/next n increments of a constant-rate load case"""
    s = copy.copy(lc)
    s.t = lc.t / lc.N * n
    s.N = n
    return s


def adjoint_cost(args):
    """Wall time of forward + adjoint gradient over a plain forward: 32^3 cells,
    50 grains, uniaxial tension 60 increments to 3 %, loss = mean squared
    deviation of the P_xx-F_xx curve from a target curve computed by the same
    solver at seeded perturbed parameters (each within +-10 %; not timed)."""
    import numpy as np
    from damask_amd import adjoint, config
    from damask_amd.grid import GridMechSpectralBasic
    grid, mat, lc = workload(32, 50, args.seed, 60, 0.03)
    num = config.Numerics()
    theta0 = adjoint.get_params(mat.phases[0])
    rng = np.random.default_rng(args.seed + 1)
    theta_t = theta0 * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, theta0.size))
    mat_t = copy.deepcopy(mat)
    mat_t.phases[0] = adjoint.set_params(mat.phases[0], theta_t)
    st_ = GridMechSpectralBasic(grid, mat_t, num)
    rt = st_.run_loadcase(lc)
    assert rt["rc"] == 0, "target run did not converge"
    target = adjoint._P33(rt["P_av"])[:, 0, 0]
    st_.close()
    s = GridMechSpectralBasic(grid, mat, num)
    t0 = time.perf_counter()
    r = s.run_loadcase(lc)
    t_fwd = time.perf_counter() - t0
    assert r["rc"] == 0
    s.close()
    s = GridMechSpectralBasic(grid, mat, num)
    s.adjoint_tape(2 * lc.N + 16)
    t0 = time.perf_counter()
    r = s.run_loadcase(lc)
    t_fwd_tape = time.perf_counter() - t0
    assert r["rc"] == 0
    info = s.adjoint_info()
    P = adjoint._P33(info["P_av"])
    n = P.shape[0]
    Pb = np.zeros_like(P)
    Pb[:, 0, 0] = 2.0 * (P[:, 0, 0] - target) / n
    t0 = time.perf_counter()
    g, st = adjoint.curve_gradient(s, 0, Pb, args.adjoint_rtol)
    t_bwd = time.perf_counter() - t0
    s.close()
    return {"workload": "32^3 FCC phenopowerlaw, 50 grains, tension 60 increments to 3 %, gradient of the "
                        "P_xx-curve MSE against the curve at seeded +-10 % parameters, w.r.t. "
                        "(xi_0, xi_inf, h_0, n, a, h_sl-sl[7])",
            "forward_s": t_fwd, "forward_with_tape_s": t_fwd_tape, "backward_s": t_bwd,
            "adjoint_over_forward": (t_fwd_tape + t_bwd) / t_fwd, "increments": n,
            "loss": float(np.mean((P[:, 0, 0] - target) ** 2)),
            "gmres_iterations_per_increment": float(st[:, 0].mean()), "gmres_rtol": args.adjoint_rtol,
            "max_gmres_relative_residual": float(st[:, 1].max()),
            "gradient_norm": float(np.linalg.norm(g))}


def progress(msg):
    """one line per finished measurement on stderr (the JSON line stays alone on stdout)"""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def batched_cost(args):
    """256 independent RVEs (edge 8-32, 10-50 grains, xi_0 and n drawn per RVE,
    tension along a random direction, 40 increments to 2 %): batched forward as
    concurrent handles and in lockstep rounds, and taped lockstep forward +
    lockstep reverse sweeps (the gradient), with the concurrent reverse sweeps
    beside them."""
    import numpy as np
    from damask_amd import batch
    items = batch.random_batch(args.batch, args.seed)
    b = batch.Batch(items)
    r = b.run(threads=args.batch_threads)
    updates = sum(s.counters()["updates"] for s in b.solvers)
    cells = int(b.cells.sum())
    b.close()
    conc = {"threads": args.batch_threads, "forward_s": r["wall_s"], "updates_per_s": updates / r["wall_s"],
            "note": "dmk_batch_run: one handle (stream) per RVE driven by a native thread pool"}
    if not args.lockstep_threads:
        return {"workload": f"{args.batch} RVEs, edge 8-32, 10-50 grains, xi_0 and n per RVE, tension along a "
                            "random direction, 40 increments to 2 %", "cells": cells, "concurrent": conc}
    b = batch.Batch(items)
    rl = b.run(threads=args.lockstep_threads, lockstep=True)
    ul = sum(s.counters()["updates"] for s in b.solvers)
    b.close()
    b = batch.Batch(items)
    b.tape()
    rt = b.run(threads=args.lockstep_threads, lockstep=True)
    curves = b.curves()
    Pb = []
    for c in curves:
        p = np.zeros_like(c)
        p[:, 0, 0] = 2.0 * (c[:, 0, 0] - 1.01 * c[:, 0, 0]) / c.shape[0]
        Pb.append(p)
    g, rb = b.gradients(Pb, rtol=args.adjoint_rtol, threads=args.lockstep_threads, lockstep=True)
    gc, rbc = b.gradients(Pb, rtol=args.adjoint_rtol, threads=args.batch_threads)
    b.close()
    conc["backward_s"] = rbc["wall_s"]
    conc["gradients_equal_lockstep"] = bool(all(np.array_equal(x, y) for x, y in zip(g, gc)))
    return {"workload": f"{args.batch} RVEs, edge 8-32, 10-50 grains, xi_0 and n per RVE, tension along a random "
                        "direction, 40 increments to 2 %; dmk_batch_run_lockstep / dmk_batch_adjoint_lockstep: "
                        "the RVEs in flight (one host thread each) run their residual evaluations, and their "
                        "reverse sweeps' adjoint operators and dot products, as rounds of one launch per kernel",
            "cells": cells, "threads": args.lockstep_threads, "forward_s": rl["wall_s"],
            "updates_per_s": ul / rl["wall_s"], "forward_with_tape_s": rt["wall_s"], "backward_s": rb["wall_s"],
            "adjoint_over_forward": (rt["wall_s"] + rb["wall_s"]) / rl["wall_s"], "gmres_rtol": args.adjoint_rtol,
            "gradient_norm": float(np.sqrt(sum(float(x @ x) for x in g))), "concurrent": conc}


def mixed_batch_cost(args):
    """The large batched configuration: 64 RVEs of 48^3 cells, 300 grains each,
    the 24-system BCC phase (Schmid law), each under its own seeded random
    symmetric deformation rate for 50 increments.  Batched forward, then a
    batched forward on a checkpointed tape (a checkpoint every `segment`
    increments, segment ~ sqrt(increments)) and the batched gradient of the
    curve loss, whose reverse sweep runs the earlier segments again; then, if
    the whole tape fits, the same without checkpoints."""
    import math

    import numpy as np
    import torch
    from damask_amd import batch
    items = batch.mixed_load_batch(args.mixed_batch, args.seed, edge=48, grains=300, increments=args.mixed_increments)
    t0 = time.perf_counter()
    b = batch.Batch(items)
    t_init = time.perf_counter() - t0
    r = b.run(threads=args.batch_threads)
    updates = sum(s.counters()["updates"] for s in b.solvers)
    cells = int(b.cells.sum())
    b.close()
    progress(f"mixed batch forward: {r['wall_s']:.1f} s")
    seg = int(math.ceil(math.sqrt(args.mixed_increments)))
    b = batch.Batch(items)
    b.tape(segment=seg)
    rt = b.run(threads=args.batch_threads)
    curves = b.curves()
    Pb = [2.0 * (c - 1.01 * c) / c.shape[0] for c in curves]
    g, rb = b.gradients(Pb, rtol=args.adjoint_rtol, threads=args.batch_threads)
    b.close()
    progress(f"mixed batch, checkpointed: forward with tapes {rt['wall_s']:.1f} s, reverse sweeps {rb['wall_s']:.1f} s")
    nsl = items[0][1].phases[0].nsl
    n = args.mixed_increments
    full_tape = (n + 1) * (27 + nsl) * 8 * cells
    seg_tape = ((seg + 1) * (27 + nsl) + (-(-n // seg) + 1) * (2 * 9 + 45 + 2 * nsl)) * 8 * cells
    # the whole tape (no recomputation) where it fits beside the solver buffers
    # and a full pool of reverse sweeps: capacity n + 4 increments (this load
    # has no cutbacks; a run that recorded more would fail)
    whole = None
    b = batch.Batch(items)
    cap = n + 4
    need = (cap + 1) * (27 + nsl) * 8 * cells
    free, total = torch.cuda.mem_get_info()
    # a sweep's work buffers: about 34 kB per cell of the 24-system phase
    work = 34_000 * int(b.cells.max())
    if need + args.batch_threads * work + total // 16 < free:
        b.tape(segment=0, capacity=cap)
        rw = b.run(threads=args.batch_threads)
        progress(f"mixed batch, whole tape: forward with tapes {rw['wall_s']:.1f} s")
        gw, rbw = b.gradients(Pb, rtol=args.adjoint_rtol, threads=args.batch_threads)
        whole = {"capacity": cap, "tape_bytes": need, "forward_with_tape_s": rw["wall_s"], "backward_s": rbw["wall_s"],
                 "adjoint_over_forward": (rw["wall_s"] + rbw["wall_s"]) / r["wall_s"],
                 "gradients_equal_checkpointed": all(np.array_equal(x, y) for x, y in zip(g, gw))}
        progress(f"mixed batch, whole tape: adjoint / forward {whole['adjoint_over_forward']:.2f}")
    b.close()
    return {"workload": f"{args.mixed_batch} RVEs of 48^3, 300 grains, BCC (cI, N_sl = [12, 12], Schmid law), "
                        f"seeded random symmetric dot_F (equivalent rate 1e-3/s, |trace| <= 1e-4/s), {n} increments",
            "cells": cells, "threads": args.batch_threads, "init_s": t_init, "forward_s": r["wall_s"],
            "updates_per_s": updates / r["wall_s"], "segment": seg, "forward_with_tape_s": rt["wall_s"],
            "backward_s": rb["wall_s"], "adjoint_over_forward": (rt["wall_s"] + rb["wall_s"]) / r["wall_s"],
            "tape_bytes_full": full_tape, "tape_bytes_checkpointed": seg_tape, "gmres_rtol": args.adjoint_rtol,
            "gradient_norm": float(np.sqrt(sum(float(x @ x) for x in g))), "whole_tape": whole}


def bcc_cost(args, n=20):
    """The headline workload with the BCC phase (cI, N_sl = [12, 12]: 24 slip
    systems, pure Schmid law): warm-up, then n timed increments"""
    from damask_amd import config, workloads
    from damask_amd.grid import GridMechSpectralBasic
    grid, mat, lc = workload(args.cells, args.grains, args.seed, args.increments, args.strain,
                             phase=workloads.BCC_PHASE)
    s = GridMechSpectralBasic(grid, mat, config.Numerics())
    if args.warmup:
        assert s.run_loadcase(sub_loadcase(lc, args.warmup))["rc"] == 0
    c0 = s.counters()
    t0 = time.perf_counter()
    r = s.run_loadcase(sub_loadcase(lc, n))
    elapsed = time.perf_counter() - t0
    assert r["rc"] == 0, "bcc: timed increments did not converge"
    c1 = s.counters()
    s.close()
    return {"workload": f"{args.cells}^3 BCC (cI, 24 slip systems, Schmid law) phenopowerlaw tension, "
                        f"{args.grains} grains, increments {args.warmup + 1}-{args.warmup + n}",
            "value": (c1["updates"] - c0["updates"]) / elapsed, "unit": "material-point updates/s",
            "ms_per_increment": elapsed / n * 1e3,
            "residual_evaluations_per_increment": (c1["residuals"] - c0["residuals"]) / n}


def schemes_cost(args):
    """The other mechanical schemes of the load file (solver: mechanical) on the
    headline workload: same warm-up and timed increments, material-point
    updates/s counted the same way (one update per cell per residual evaluation)."""
    from damask_amd import config
    from damask_amd.grid import grid_mechanical
    grid, mat, lc = workload(args.cells, args.grains, args.seed, args.increments, args.strain)
    out = {}
    for name in ("spectral_polarization", "spectral_Galerkin"):
        s = grid_mechanical(name)(grid, mat, config.Numerics())
        if args.warmup:
            assert s.run_loadcase(sub_loadcase(lc, args.warmup))["rc"] == 0
        c0 = s.counters()
        t0 = time.perf_counter()
        r = s.run_loadcase(sub_loadcase(lc, args.steps))
        elapsed = time.perf_counter() - t0
        assert r["rc"] == 0, f"{name}: timed increments did not converge"
        c1 = s.counters()
        its = r["stats"][:, 2]
        out[name] = {"value": (c1["updates"] - c0["updates"]) / elapsed, "unit": "material-point updates/s",
                     "ms_per_increment": elapsed / args.steps * 1e3,
                     "residual_evaluations_per_increment": (c1["residuals"] - c0["residuals"]) / args.steps,
                     "iterations_per_increment": float(its.mean()),
                     "P_xx_last_increment": float(r["P_av"][-1, 0])}
        s.close()
    return out


def kernel_breakdown(args, grid, mat, lc, n=10):
    """HIP-event time per kernel group and increment over increments after the
    warm-up, in a separate run with every group timed (not the timed region)"""
    from damask_amd import config
    from damask_amd.grid import GridMechSpectralBasic
    s = GridMechSpectralBasic(grid, mat, config.Numerics())
    if args.warmup:
        assert s.run_loadcase(sub_loadcase(lc, args.warmup))["rc"] == 0
    s.profile(2)
    t0 = time.perf_counter()
    assert s.run_loadcase(sub_loadcase(lc, n))["rc"] == 0
    wall = time.perf_counter() - t0
    kt = s.kernel_times()
    s.close()
    out = {k: round(v[0] / n, 4) for k, v in kt.items()}
    out["wall_with_all_events"] = round(wall * 1e3 / n, 4)
    out["increments"] = f"{args.warmup + 1}-{args.warmup + n}"
    return out, kt


def spectral_roofline(kt, cells):
    """Achieved HBM bandwidth of the residual's transform path (HIP events per
    kernel group): algorithmic bytes (each field read and written once per
    group, flops.spectral_bytes) x launches / time, against 8 TB/s"""
    from damask_amd import flops
    N, Nk = cells**3, (cells // 2 + 1) * cells**2
    bts = flops.spectral_bytes(N, Nk)
    names = {"fft_fwd": "x-y forward passes (fft_xy_r2c)",
             "gamma_div": "z passes + Gamma convolution + divergence (fft_zgz, or z passes and gamma_div_kernel)",
             "fft_inv": "y-x inverse passes (fft_yx_c2r)"}
    out = {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "kernels": {}}
    for k in ("fft_fwd", "gamma_div", "fft_inv"):
        ms, n = kt[k]
        gbs = bts[k] * n / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out["kernels"][k] = {"kernels": names[k], "bytes_per_launch": bts[k], "avg_launch_us": ms * 1e3 / max(n, 1),
                             "achieved": gbs, "frac": gbs / PEAK_HBM_GBS}
    return out


def host_cpu():
    """model, sockets and cores of this host (lscpu) and the CPUs this process may use"""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    def num(k):
        try:
            return int(info.get(k, ""))
        except ValueError:
            return None
    return {"model": info.get("Model name"), "sockets": num("Socket(s)"), "cores_per_socket": num("Core(s) per socket"),
            "threads_per_core": num("Thread(s) per core"), "logical_cpus": num("CPU(s)"),
            "cpus_available": avail}


def cpu_baseline(args):
    """The oracle (C/OpenMP restatement of the same algorithm: same NGMRES, same
    C_ref rule, same update counter) on the headline grid, in the mode that
    forms dPdF, the orientation and the stiffness averages once per solve as
    the GPU path does.  It warms up over the bench's own warm-up increments and
    times the next --cpu-increments, a slice of the GPU's timed increments.
    The FFT share is the time of the oracle's transforms (two r2c and one c2r
    of the nine components per residual evaluation) over the timed wall time."""
    import ctypes as C

    import numpy as np
    from damask_amd import config
    from oracle import pyoracle
    cpu = host_cpu()
    threads = min(16, cpu["cpus_available"])
    if cpu["threads_per_core"] and cpu["threads_per_core"] > 1:
        threads = max(1, min(threads, cpu["cpus_available"] // cpu["threads_per_core"]))
    grid, mat, lc = workload(args.cpu_cells, args.grains, args.seed, args.increments, args.strain)
    pr = pyoracle.Problem(grid, mat, config.Numerics(), threads=threads, tangent_once=True)
    if args.warmup:
        assert pr.run_loadcase(sub_loadcase(lc, args.warmup))["rc"] == 0, "cpu baseline: warm-up did not converge"
    t0 = time.perf_counter()
    r = pr.run_loadcase(sub_loadcase(lc, args.cpu_increments))
    dt = time.perf_counter() - t0
    assert r["rc"] == 0, "cpu baseline: timed increments did not converge"
    residuals = int(r["stats"][:, 3].sum())
    # the oracle's transforms of the nine components, timed alone
    L = pyoracle.lib()
    N = grid.n_cells
    Nk = (args.cpu_cells // 2 + 1) * args.cpu_cells ** 2
    x = np.random.default_rng(0).random(9 * N)
    X = np.zeros(2 * 9 * Nk)
    t1 = time.perf_counter()
    for _ in range(3):
        L.orc_fft_r2c(pr.grid, 9, pyoracle.ptr(x), pyoracle.ptr(X))
    t_r2c = (time.perf_counter() - t1) / 3
    t1 = time.perf_counter()
    for _ in range(3):
        L.orc_fft_c2r(pr.grid, 9, pyoracle.ptr(X), pyoracle.ptr(x))
    t_c2r = (time.perf_counter() - t1) / 3
    pr.close()
    fft_share = residuals * (2 * t_r2c + t_c2r) / dt
    first = args.warmup + 1
    return {"value": r["n_updates"] / dt, "unit": "material-point updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C/OpenMP restatement, dPdF and stiffness averages once per solve as on the GPU), "
                      f"{args.cpu_cells}^3 cells, {args.grains} grains, increments {first}-"
                      f"{args.warmup + args.cpu_increments} of the same tension after {args.warmup} warm-up "
                      f"increments ({r['n_updates']} updates in {dt:.1f} s, {residuals} residual evaluations)",
            "host": cpu, "openmp_threads": threads, "fft_share": fft_share,
            "residual_evaluations_per_increment": residuals / args.cpu_increments}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box has one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.only_adjoint:
        print(json.dumps({"adjoint": adjoint_cost(args)}))
        return
    if args.only_batch:
        print(json.dumps({"batched": batched_cost(args)}))
        return
    if args.only_schemes:
        print(json.dumps({"schemes": schemes_cost(args)}))
        return
    if args.only_mixed:
        print(json.dumps({"mixed_batch": mixed_batch_cost(args)}))
        return
    import numpy as np
    import torch
    from damask_amd import config, flops
    from damask_amd.grid import GridMechSpectralBasic

    assert torch.cuda.is_available(), "bench.py needs the GPU"
    grid, mat, lc = workload(args.cells, args.grains, args.seed, args.increments, args.strain)
    t_init = time.perf_counter()
    s = GridMechSpectralBasic(grid, mat, config.Numerics())
    t_init = time.perf_counter() - t_init
    if args.warmup:
        r = s.run_loadcase(sub_loadcase(lc, args.warmup))
        assert r["rc"] == 0, "warm-up did not converge"
    c0 = s.counters()
    fft_path = [int(v) for v in s.scalar("fft_path")]
    # HIP events around the stress kernel only: an event record between two
    # kernels leaves the stream idle for a few microseconds
    s.profile(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = s.run_loadcase(sub_loadcase(lc, args.steps))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert r["rc"] == 0, "timed increments did not converge"
    c1 = s.counters()
    kt = s.kernel_times()
    s.profile(0)
    s.close()
    breakdown, kt_all = kernel_breakdown(args, grid, mat, lc)
    updates = c1["updates"] - c0["updates"]
    newton = c1["newton"] - c0["newton"]
    state = c1["state"] - c0["state"]
    residuals = c1["residuals"] - c0["residuals"]
    value = updates / elapsed
    nsl = mat.phases[0].nsl
    stress_ms, stress_n = kt["stress"]
    fl = flops.stress_kernel_flops(nsl, newton, state)
    achieved = fl / (stress_ms * 1e-3) / 1e12
    line = {
        "metric": "material-point updates/s (64^3 FCC phenopowerlaw, FP64)",
        "value": value,
        "unit": "material-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Voronoi polycrystal and orientations)",
        "config": {"workload": f"{args.cells}^3 FCC phenopowerlaw uniaxial tension, {args.grains} grains, "
                               f"{args.increments} increments to {args.strain * 100:g} % (increments "
                               f"{args.warmup + 1}-{args.warmup + args.steps} timed)",
                   "cells": args.cells, "grains": args.grains, "solver": "spectral_basic + NGMRES",
                   "residual_evaluations": residuals,
                   "residual_evaluations_per_increment": residuals / args.steps,
                   "newton_iterations_per_update": newton / max(updates, 1),
                   "state_iterations_per_update": state / max(updates, 1),
                   "init_s": t_init,
                   "fft_path": {"xy_threads": fft_path[0], "xy_in_place": fft_path[1],
                                "zgz_columns": fft_path[2], "zgz_threads": fft_path[3],
                                "zgz_in_place": fft_path[4]},
                   "stress_kernel_ms": round(kt["stress"][0], 3),
                   "kernel_ms_per_increment": breakdown,
                   "cells_x_increments_per_s": args.cells**3 * args.steps / elapsed},
        "roofline": {"bound": "fp64-valu", "compute": "fp64 vector ALU, no MFMA on the path (MI355X fp64 vector "
                                                      "peak 78.6 TFLOP/s)",
                     "kernel": "stress_kernel<12, true>", "achieved": achieved,
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": args.traffic_bytes or None,
                     "algorithmic_bytes_per_launch": flops.stress_kernel_bytes(nsl) * args.cells**3,
                     "flops_per_launch": fl / max(stress_n, 1),
                     "avg_launch_ms": stress_ms / max(stress_n, 1),
                     # measured HBM bytes per launch over the launch time, against 8 TB/s
                     "traffic_GBs": (args.traffic_bytes / (stress_ms / max(stress_n, 1) * 1e-3) / 1e9
                                     if args.traffic_bytes else None),
                     "traffic_frac_of_hbm_peak": (args.traffic_bytes / (stress_ms / max(stress_n, 1) * 1e-3) / 1e9
                                                  / PEAK_HBM_GBS if args.traffic_bytes else None)},
    }
    line["roofline_spectral"] = spectral_roofline(kt_all, args.cells)
    progress(f"forward: {value / 1e6:.1f} M updates/s")
    if args.adjoint:
        line["adjoint"] = adjoint_cost(args)
        progress(f"adjoint / forward: {line['adjoint']['adjoint_over_forward']:.2f}")
    if args.batch:
        line["batched"] = batched_cost(args)
        bt = line["batched"]
        if "updates_per_s" in bt:
            progress(f"batched: {bt['updates_per_s'] / 1e6:.1f} M updates/s in lockstep "
                     f"({bt['concurrent']['updates_per_s'] / 1e6:.1f} M concurrent), adjoint / forward "
                     f"{bt['adjoint_over_forward']:.2f}")
        else:
            progress(f"batched: {bt['concurrent']['updates_per_s'] / 1e6:.1f} M updates/s concurrent")
    if args.schemes:
        line["schemes"] = schemes_cost(args)
        progress("other schemes timed")
    if args.mixed_batch:
        line["mixed_batch"] = mixed_batch_cost(args)
        progress(f"mixed batch: {line['mixed_batch']['updates_per_s'] / 1e6:.1f} M updates/s, adjoint / forward "
                 f"{line['mixed_batch']['adjoint_over_forward']:.2f}")
    if args.bcc:
        line["bcc"] = bcc_cost(args)
        progress(f"bcc: {line['bcc']['value'] / 1e6:.1f} M updates/s")
    if args.cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line))


if __name__ == "__main__":
    main()
