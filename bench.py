#!/usr/bin/env python3
"""Benchmark of the wPFAW3D per-time-step hot path on one MI355X.

Workload: the 256^3 wormholing configuration (16,777,216 cells, seeded
Gaussian-random-field rock, inlet plane at 1e5 Pa, outlet plane at "p = 0"),
run for its full 500 steps.  A step is one pass of the hot path: dt rule,
permeability, pressure solve, velocity + phase-field + concentration + Euler
update, with every input already resident in HBM.

Two paths run the same 500 steps from the same initial state:
  plain   FP64 MG-preconditioned CG, rtol 1e-10       (the FP64 reference path)
  fast    FP32 MG-PCG inside FP64 iterative refinement, rtol 1e-10 on the
          FP64 true residual (the headline path)
Every step of both runs is timed.  `value` is the fast path's time-steps/s
over the timed window, steps W+1..W+K (--warmup W, --steps K; by default the
whole run, W = 0 and K = 500); `whole_run` gives both paths' rates over all
500 steps and over each 100-step window, so the figure does not depend on the
window chosen.  The fast path's relative pressure error is reported against
the plain path after the run, and against the FP64 oracle for one step at
this size (both GPU paths and the oracle's BiCGStab + SOR step from the
state the timed window starts from).

roofline: the dominant kernel of the fast path (by default the CG update
fused with the first pre-smoothing step, k_vupsweep: the largest share of a
step's kernel time), timed with HIP events on the library's stream over the
timed window; achieved = algorithmic bytes per launch / average launch
time; traffic = HBM bytes per launch measured with rocprofv3 counters.
cpu_baseline: the FP64 oracle (C restatement, OpenMP) on this host, from the
state the timed window starts from: one step with the reference's own solver
(BiCGStab + SOR, 10 iterations, omega 1.7, one SOR block per thread as the
ranks of an mpirun, rtol 1e-10, maxit 1000) is `value`; one step with CG +
Jacobi (the CPU configuration's solver, same rtol and cap) is reported beside
it, with whether each solve converged.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-steps/s at 256³: learned+FP32-refined vs plain FP64 GPU, ≤1e-6 p err"
HBM_PEAK_GBS = 8000.0
# what a kernel that streams three or more 256 MiB arrays reaches on this
# part (the streaming probe, 5.39-5.62 TB/s: profiles/streaming_bandwidth.md)
STREAM_GBS = 5500.0

# algorithmic bytes per cell (each array element the math must read or
# write once per launch; mask = 1 byte); r = vector precision, h = stored
# face precision of the fine operator
def smoother_bytes(r: int, h: int, cheb: bool, launches: int, its: int) -> float:
    """Mean over the fine-level smoothing launches of the timed steps.  With
    two sweeps and the first pre-smoothing step fused into the CG update,
    each V-cycle launches
      post step 1   x, b, invD, 3 faces, mask, coarse correction -> y   4r + 3h + 1 + r/8
      post step 2   + x0 (Chebyshev; Jacobi needs none)                 5r + 3h + 1 + r/8
    and the first V-cycle of every inner solve adds a pre-smoothing step
    x, b, invD, 3 faces -> y (4r + 3h).  Inner solves = launches - 2 its."""
    post1 = 4 * r + 3 * h + 1 + r / 8
    post2 = post1 + (r if cheb else 0)
    pre = max(0, launches - 2 * its)
    return (its * (post1 + post2) + pre * (4 * r + 3 * h)) / max(1, launches)


def bytes_per_cell(kernel: str, r: int, h: int, cheb: bool = True, launches: int = 1,
                   its: int = 1) -> float:
    return {
        "smooth": smoother_bytes(r, h, cheb, launches, its),
        # z, p_old, 3 faces, mask -> p_new, q (the CG operator: with FP16
        # preconditioner faces, h = 2, the matvec still reads FP32 faces)
        "matvec": 4 * r + 3 * max(h, r) + 1,
        # CG update fused with the first pre-smoothing step:
        # e, p, r, q, 3 faces, mask -> e, r', y (1/D formed from the faces)
        "cg_update": 7 * r + 3 * h + 1,
        "restrict": 2 * r + 3 * h + 1 + r / 8,
        "true_resid": 5 * 8 + 1 + r,  # x, f, t x3 (FP64), mask -> r (precision r); 1/D from the faces
        "step_fields": 6 * 8 + 1 + 2 * 8,  # phi, c, p, t x3, mask -> phi', c'
        "perm": 2 * 8 + 1 + 4 * 8,  # phi, krock, mask -> tx, ty, tz, invD
    }[kernel]


# HBM traffic per launch of the roofline kernel, from two rocprofv3 --pmc
# runs of this benchmark (read requests by size, then write requests by
# size; see traffic_table), written by --traffic-table (the commands and the
# per-kernel table are in the same folder)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_256.json")
READ_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")
WRITE_COUNTERS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")
# kernels of the fast path reported in the table: name fragment -> (key,
# algorithmic bytes per cell with FP32 vectors, the preconditioner's FP16
# faces and the CG operator's FP32 faces: the formulas of bytes_per_cell;
# post1/post2 are smoother_bytes' two post-smoothing launches; Gram = 4
# changes, r and the FP32 faces of their stencil products; projection update
# = x (FP64), r, 4 changes, FP32 faces, invD, mask -> x, r)
TRAFFIC_KERNELS = {
    # the update form only (LDS-staged, UPD); the first V-cycle's form
    # without the update is another instance
    "k_vupsweepIffDF16_Li4ELb1ELb1E": ("cg_update", bytes_per_cell("cg_update", 4, 2)),
    # the sweeps form 1/D from the faces: b, x, xo, 3 faces, mask, xc -> y
    "k_vsweepIfDF16_Li4ELb1ELb0ELb1E": ("smooth_post2", 4 * 4 + 3 * 2 + 1 + 4 / 8),
    "k_vsweepIfDF16_Li4ELb0ELb1ELb0E": ("smooth_post1", 3 * 4 + 3 * 2 + 1 + 4 / 8),
    "k_vdir_mv<float, float, 4>": ("matvec", bytes_per_cell("matvec", 4, 2)),
    "k_vfresid_restrictIfDF16": ("restrict", bytes_per_cell("restrict", 4, 2)),
    "k_step_fields<double>": ("step_fields", bytes_per_cell("step_fields", 4, 2)),
    # the fast path's k_perm also writes the FP32 faces, their FP16 copies
    # and an FP32 invD
    "k_perm<2>": ("perm", bytes_per_cell("perm", 4, 2) + 3 * 4 + 3 * 2 + 4),
    "k_vtrue_resid<2>": ("true_resid", bytes_per_cell("true_resid", 4, 2)),
    # the 5-deep projection ring: 5 vectors, r, 3 FP32 faces; x (FP64) in
    # and out, 5 vectors, 3 faces, r in and out, the updated r's invD, mask
    "k_vgram<float, float, 4, 5>": ("gram", 4 * 5 + 4 + 3 * 4),
    "k_vproj_update<float, float, 4, 5": ("proj_update", 8 * 2 + 4 * 5 + 3 * 4 + 4 * 2 + 4 + 1),
    "k_coarsen_pairsIfDF16": ("coarsen_pairs", 9.5),
}


def traffic_table(dir_read, dir_write, grid=256):
    """HBM bytes per launch of each kernel from the counter CSVs of two
    rocprofv3 --pmc runs (READ_COUNTERS, then WRITE_COUNTERS): read bytes =
    32 n32 + 64 n64 + 128 n128 and write bytes = 32 (n - n64) + 64 n64 over
    the TCC/EA request counts by size.  (On this part FETCH_SIZE counts
    128-byte reads through TCC_BUBBLE, which reads 0, so FETCH_SIZE
    undercounts them; the request-size counters reproduce the stream
    probes' known bytes exactly.)  Prints a markdown table and writes
    TRAFFIC_FILE."""
    import csv
    from collections import defaultdict

    def load(d):
        per = defaultdict(lambda: defaultdict(float))
        n = defaultdict(set)
        with open(os.path.join(d, "run_counter_collection.csv")) as f:
            for r in csv.DictReader(f):
                key = next((v[0] for k, v in TRAFFIC_KERNELS.items() if k in r["Kernel_Name"]), None)
                if key is None:
                    continue
                per[key][r["Counter_Name"]] += float(r["Counter_Value"])
                n[key].add(r["Dispatch_Id"])
        return per, {k: len(v) for k, v in n.items()}

    rd, nr = load(dir_read)
    wr, nw = load(dir_write)
    cells = grid ** 3
    out = {"grid": grid, "kernels": {}}
    print("| kernel | launches | read MB | written MB | B/cell | algorithmic B/cell | ratio |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for key, alg in TRAFFIC_KERNELS.values():
        if key not in rd or key not in wr:
            continue
        a, b = rd[key], wr[key]
        rbytes = (32 * a["TCC_EA0_RDREQ_32B_sum"] + 64 * a["TCC_EA0_RDREQ_64B_sum"]
                  + 128 * a["TCC_EA0_RDREQ_128B_sum"]) / nr[key]
        wbytes = (32 * (b["TCC_EA0_WRREQ_sum"] - b["TCC_EA0_WRREQ_64B_sum"])
                  + 64 * b["TCC_EA0_WRREQ_64B_sum"]) / nw[key]
        out["kernels"][key] = {"bytes_per_launch": rbytes + wbytes, "read": rbytes, "written": wbytes,
                               "launches": nr[key], "algorithmic_bytes_per_cell": alg}
        per_cell = (rbytes + wbytes) / cells
        print(f"| {key} | {nr[key]} | {rbytes / 1e6:.1f} | {wbytes / 1e6:.1f} | {per_cell:.2f} | {alg:.2f} "
              f"| {per_cell / alg:.3f} |")
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(out, f, indent=1)


def measured_traffic(kernel: str, grid: int):
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if t.get("grid") != grid or not k:
        return None
    return k["bytes_per_launch"]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def run_path(case, opts, warmup, steps, run_steps, profile_kernel=None, net=None):
    """Runs max(run_steps, warmup + steps) steps; steps warmup+1..warmup+steps
    form the timed window (HIP-event profile of one kernel over it).  Every
    step is timed on the host (wpf_step ends on a stream synchronisation)."""
    from mizuho_edit_amd.wpfaw3d import WPFAW3D
    g = WPFAW3D(case, device=0, opts=opts)
    g.set_tile_depth(-1)  # autotuned tile depth of the vector solver kernels
    if net is not None:
        g.attach_net(net)
    total = max(run_steps, warmup + steps)
    step_s, its_window, stage_steps, prof, state = [], 0, [], None, None
    wall = 0.0
    for s in range(total):
        if s == warmup:
            state = {n: g.download(n) for n in ("pressure", "phase", "concent")}
            state.update(itime=g.itime, time=g.time, dtime=g.dtime, taumin=g.taumin)
            sync()
            if profile_kernel:
                g.profile_start(profile_kernel)
            tw = time.perf_counter()
        t0 = time.perf_counter()
        rec = g.step()
        step_s.append(time.perf_counter() - t0)
        if rec["learned"]:
            stage_steps.append(s + 1)
        if warmup <= s < warmup + steps:
            its_window += rec["iterations"]
        if s == warmup + steps - 1:
            sync()
            wall = time.perf_counter() - tw
            if profile_kernel:
                prof = g.profile_stop()
        if (s + 1) % 50 == 0 or s + 1 == total:
            log(f"  step {s + 1}/{total}: its={rec['iterations']} outer={rec['outer']} "
                f"dt={rec['dtime']:.3e} ({sum(step_s[-50:]) / len(step_s[-50:]) * 1e3:.2f} ms/step)")
    g.stage_steps = stage_steps
    return g, wall, its_window, state, prof, step_s


def window_rates(step_s, width=100):
    return [{"steps": f"{a + 1}..{min(a + width, len(step_s))}",
             "steps_per_s": len(step_s[a:a + width]) / sum(step_s[a:a + width])}
            for a in range(0, len(step_s), width)]


def _oracle_from(case, state):
    from oracle.oracle import OracleState
    st = OracleState(case)
    st.pressure[...] = state["pressure"]
    st.phase[...] = state["phase"]
    st.concent[...] = state["concent"]
    st.itime, st.time, st.dtime, st.taumin = (state["itime"], state["time"], state["dtime"],
                                             state["taumin"])
    return st


def cpu_baseline(case, state, nthreads):
    """One step of each CPU solver from `state`; returns the records and the
    oracle state after the BiCGStab + SOR step."""
    out = {}
    st = _oracle_from(case, state)
    t0 = time.perf_counter()
    rec = st.step_bcgs(rtol=1e-10, maxit=1000, nblocks=nthreads)
    out["bicgstab_sor"] = dict(seconds=time.perf_counter() - t0, iterations=rec["iterations"],
                               converged=bool(rec["converged"]))
    ref = st
    st = _oracle_from(case, state)
    t0 = time.perf_counter()
    st.run(1, rtol=1e-10, maxit=1000, nthreads=nthreads)
    its = int(st.hist[-1][2])
    out["cg_jacobi"] = dict(seconds=time.perf_counter() - t0, iterations=its,
                            converged=bool(st.hist[-1][3] <= 1e-10))
    return out, ref


def gpu_one_step(case, state, opts):
    """One step of a GPU path on a fresh context from `state` (no solution
    history: the extrapolated and projected guesses start empty)."""
    from mizuho_edit_amd.wpfaw3d import WPFAW3D
    g = WPFAW3D(case, device=0, opts=opts)
    for n in ("pressure", "phase", "concent"):
        g.upload(n, state[n])
    g.itime, g.time, g.dtime, g.taumin = state["itime"], state["time"], state["dtime"], state["taumin"]
    rec = g.step()
    fields = {n: g.download(n) for n in ("pressure", "phase", "concent")}
    g.close()
    return rec, fields


def calibrate_copy(reps=20):
    """A plain 1 GiB device-to-device copy (torch's copy kernel), timed with
    events: a reference point for HBM bandwidth beside the streaming probe."""
    import torch
    n = (1 << 30) // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        y.copy_(x)
    b.record()
    b.synchronize()
    ms = a.elapsed_time(b) / reps
    print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30,
                      "us_per_copy": 1e3 * ms, "GBps_read_plus_write": 2 * (1 << 30) / (ms * 1e-3) / 1e9}),
          flush=True)


def calibrate_streams(sizes=(1 << 26, 1 << 24), reps=20):
    """The library's streaming probe (wpf_stream_probe): one JSON line per
    (arrays read, arrays written) pair and array size (n floats per array;
    256 MiB arrays keep every configuration far beyond the 256 MB
    Infinity Cache, 64 MiB ones are the size of a 256^3 FP32 field)."""
    import ctypes
    from mizuho_edit_amd import _lib
    L = _lib.lib()
    for n in sizes:
        for nr, nw in ((1, 1), (2, 1), (4, 1), (5, 2), (6, 3), (7, 1), (8, 1), (8, 3), (10, 1), (0, 3)):
            us = ctypes.c_double()
            _lib.check(L.wpf_stream_probe(0, nr, nw, n, reps, ctypes.byref(us)), "wpf_stream_probe")
            nbytes = 4.0 * n * (nr + nw)
            print(json.dumps({"read_arrays": nr, "written_arrays": nw, "mib_per_array": 4 * n >> 20,
                              "us": us.value, "GBps": nbytes / (us.value * 1e-6) / 1e9}), flush=True)


def bench_network(a):
    """The learned part's kernels alone: forward passes and energy-loss
    gradients of the shipped flexible-CG weights on an a.grid^3 input (the
    network runs at a.grid/4) with seeded random fields, a.warmup untimed and
    a.steps timed passes of each.  Prints one JSON line with host-timed
    milliseconds per pass (the host copies of the fields included);
    rocprofv3 gives the per-kernel times and MFMA counters."""
    import ctypes
    import numpy as np
    from mizuho_edit_amd import _lib, learned
    net = learned.PCNet()
    net.load(learned.WEIGHTS_FCG)
    n = a.grid
    L = _lib.lib()
    rng = np.random.default_rng(a.seed)
    # the C ABI's layouts directly (planar [3][n^3], x fastest): no host
    # reordering inside the timed passes
    fields = rng.standard_normal((3, n ** 3), dtype=np.float32)
    target = rng.standard_normal(n ** 3, dtype=np.float32)
    faces = rng.uniform(0.5, 1.5, (3, n, n, n)).astype(np.float32)  # [axis][z][y][x]
    faces[0, :, :, -1] = 0.0  # no face beyond the global edge
    faces[1, :, -1, :] = 0.0
    faces[2, -1, :, :] = 0.0
    faces = faces.reshape(3, -1)
    fixed = np.zeros((n, n, n), np.uint8)
    fixed[:, :, 0] = fixed[:, :, -1] = 1  # Dirichlet planes x = 0 and x = n-1
    fixed = fixed.reshape(-1)
    y = np.zeros((n // 4) ** 3, np.float32)
    grad = np.zeros(net.nparams, np.float32)
    loss = ctypes.c_double()
    fp = lambda v: v.ctypes.data_as(_lib.F)  # noqa: E731

    def forward():
        _lib.check(L.wpf_net_forward(net._h, n, n, n, fp(fields), fp(y)), "wpf_net_forward")

    def energy_gradient():
        _lib.check(L.wpf_net_gradient_energy(net._h, n, n, n, fp(fields), fp(target), fp(faces),
                                             fixed.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                             ctypes.byref(loss), fp(grad)), "wpf_net_gradient_energy")

    out = {"metric": "learned-part pass time", "unit": "ms", "grid": n, "network_grid": n // 4,
           "params": int(net.nparams), "weights": os.path.relpath(learned.WEIGHTS_FCG, ROOT),
           "steps": a.steps, "warmup": a.warmup, "data": "synthetic: seeded random fields"}
    for name, fn in (("forward", forward), ("energy_gradient", energy_gradient)):
        for _ in range(a.warmup):
            fn()
        sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            fn()
        sync()
        out[f"{name}_ms"] = 1e3 * (time.perf_counter() - t0) / a.steps
        log(f"{name}: {out[f'{name}_ms']:.2f} ms per pass")
    net.close()
    print(json.dumps(out), flush=True)


def bench_learned(a):
    """The learned configuration: the 192^3 workload (a realisation the
    shipped weights were not trained on) for a.learned steps, solved with the
    reference's CG + Jacobi (rtol 1e-10, maxit 1000, from the previous
    pressure) and with flexible CG carrying the learned stage (the shipped
    energy-trained weights, the stage at every 16th preconditioner
    application, plain CG continuing from the iterate when a flexible pass
    stalls).  Both runs start from the same state; every step is timed.
    Prints one JSON line: iterations, solve and wall seconds, the steps that
    hit the 1000-iteration cap, stage applications, per 100 steps, and how
    far the two runs' final fields are apart."""
    import numpy as np
    from mizuho_edit_amd import _lib, io, learned
    from mizuho_edit_amd.wpfaw3d import WPFAW3D, solver_opts
    n, nsteps = 192, a.learned
    case = io.workload("learned-192", seed=1001, grid=(n, n, n))
    weights = a.weights or learned.WEIGHTS_FCG
    net = learned.PCNet()
    net.load(weights)
    runs = {}
    for name, use in (("cg_jacobi", False), ("flexible_cg_learned", True)):
        log(f"{name}: {nsteps} steps at {n}^3")
        g = WPFAW3D(case, device=0)
        flags = 0
        if use:
            g.attach_net(net)
            flags = _lib.FLAG_LEARNED_FCG
        o = solver_opts("cg", rtol=1e-10, maxit=1000, flags=flags)
        its, solve_s, capped, stages, step_s = [], 0.0, 0, 0, []
        for s in range(nsteps):
            t0 = time.perf_counter()
            rec = g.step(o)
            step_s.append(time.perf_counter() - t0)
            its.append(rec["iterations"])
            solve_s += rec["seconds"]
            capped += 0 if rec["converged"] else 1
            stages += rec["learned"]
            if (s + 1) % 50 == 0:
                log(f"  step {s + 1}/{nsteps}: its={rec['iterations']} "
                    f"({sum(its[-50:]) / 50:.0f} per step over the last 50)")
        runs[name] = {"iterations": int(sum(its)), "iterations_per_step": sum(its) / nsteps,
                      "solve_seconds": solve_s, "wall_seconds": sum(step_s),
                      "steps_at_cap": capped, "stage_applications": stages,
                      "iterations_per_100_steps": [int(sum(its[k:k + 100])) for k in range(0, nsteps, 100)],
                      "fields": {f: g.download(f) for f in ("pressure", "phase", "concent")}}
        g.close()
    net.close()
    a0, a1 = runs["cg_jacobi"], runs["flexible_cg_learned"]
    diff = {f: float(np.abs(a1["fields"][f] - a0["fields"][f]).max() / np.abs(a0["fields"][f]).max())
            for f in ("pressure", "phase", "concent")}
    for r in runs.values():
        del r["fields"]
    out = {"metric": "learned configuration: iterations and solve time, flexible CG with the learned "
                     "stage against CG + Jacobi", "grid": [n, n, n], "steps": nsteps,
           "weights": os.path.relpath(weights, ROOT),
           "data": "synthetic: seeded GRF rock (seed 1001, not in the training set)",
           "runs": runs,
           "iteration_ratio": a1["iterations"] / a0["iterations"],
           "solve_time_ratio": a1["solve_seconds"] / a0["solve_seconds"],
           "final_field_rel_diff": diff}
    print(json.dumps(out), flush=True)


def bench_ensemble(a):
    """The ensemble workload (48 uneven grids, 40,894,464 cells): the same
    steps once member by member with wpf_step (per-instance launches, one
    host round trip per CG iteration) and once as one ensemble (batched
    launches over all members, captured graphs).  Prints one JSON line."""
    import numpy as np
    from mizuho_edit_amd import io
    from mizuho_edit_amd._lib import FLAG_CHEBYSHEV, FLAG_EXTRAPOLATE, FLAG_OP16, FLAG_PROJECT
    from mizuho_edit_amd.ensemble import Ensemble
    from mizuho_edit_amd.wpfaw3d import WPFAW3D, solver_opts
    log(f"building the ensemble workload (shrink {a.ens_shrink})")
    cases = io.ensemble_cases(shrink=a.ens_shrink)
    cells = sum(c.params.nx * c.params.ny * c.params.nz for c in cases)
    opts = solver_opts("mixed", rtol=1e-10, inner_rtol=1e-4,
                       flags=FLAG_EXTRAPOLATE | FLAG_CHEBYSHEV | FLAG_OP16 | (0 if a.no_project else FLAG_PROJECT))
    res = {}
    # the ensemble takes the last warm-up steps itself, so that its adaptive
    # schedule has settled before the timed steps (the same steps both ways)
    settle = min(10, a.warmup)
    for mode in ("per_instance", "ensemble"):
        grids = [WPFAW3D(c, device=0, opts=opts) for c in cases]
        for g in grids:
            g.run(a.warmup - (settle if mode == "ensemble" else 0))
        sync()
        if mode == "per_instance":
            t0 = time.perf_counter()
            for s in range(a.steps):
                for g in grids:
                    g.step()
                log(f"  per-instance step {s + 1}/{a.steps}")
            sync()
            wall = time.perf_counter() - t0
            info = {}
        else:
            ens = Ensemble(grids, opts)
            tg, tz = ens.tune()
            s0 = ens.step(settle)  # captures the graphs, adapts the schedule
            log(f"  ensemble tuned: {tg} workgroups, {tz} planes per tile; schedule "
                f"{s0['stages']} stages, {s0['inner_iters']} iteration slots")
            st = ens.step(a.steps)
            wall = st["seconds"]
            ens.sync()
            info = {"grid_blocks": tg, "tile_depth": tz, "fallbacks": st["fallbacks"],
                    "schedule": f"{st['stages']} stages, {st['inner_iters']} FP32 iteration slots in all (adaptive)",
                    "fp32_iterations_per_member_step": st["iterations"] / (a.steps * len(grids))}
            ens.close()
        res[mode] = (wall, info)
        for g in grids:
            g.close()
    wall_i, _ = res["per_instance"]
    wall_e, info = res["ensemble"]
    out = {
        "metric": "aggregate cell-updates/s, ensemble of 48 uneven grids",
        "value": cells * a.steps / wall_e,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "higher_is_better": True,
        "per_instance_value": cells * a.steps / wall_i,
        "speedup_vs_per_instance": wall_i / wall_e,
        "ms_per_step": 1e3 * wall_e / a.steps,
        "dtype": "f64 stencils + f32 solve (FP16-stored faces) in f64 refinement",
        "data": "synthetic: 48 seeded GRF rocks, correlation length 2-8 cells, mean porosity 0.12-0.25",
        "learned_stage": "not used: the ensemble's captured schedule holds no per-member network "
                         "launches (the mixed-precision path without the network)",
        "config": {"workload": "ensemble: 16x64^3 + 16x96x64x64 + 8x128x96x96 + 8x160x128x128"
                               + (f" (each dimension / {a.ens_shrink})" if a.ens_shrink > 1 else ""),
                   "cells": cells,
                   "solver": ("" if a.no_project else "projected initial guess (last 5 solution changes) + ")
                   + "FP32 MG-PCG (Chebyshev) in FP64 refinement, rtol 1e-10",
                   **info},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--run-steps", type=int, default=500,
                    help="length of the simulated run (the timed window lies inside it)")
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--kernel", default="cg_update", help="kernel reported in roofline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--weights", default=None,
                    help="network weights (.npz); default: the package's trained weights")
    ap.add_argument("--no-learned", action="store_true", help="fast path without the network")
    ap.add_argument("--jacobi", action="store_true",
                    help="fast path with the plain path's smoother and FP32 faces")
    ap.add_argument("--no-project", action="store_true",
                    help="fast path without the projected initial guess")
    ap.add_argument("--fast-levels", type=int, default=6,
                    help="multigrid levels of the fast path (the plain path keeps 5)")
    ap.add_argument("--f32-stencils", action="store_true",
                    help="fast path with FP32 phase/concentration stencils (FP64 fields and "
                         "update); slower at 256^3 since the less accurate fields cost "
                         "solver iterations")
    ap.add_argument("--calibrate-streams", action="store_true",
                    help="streaming bandwidth of kernels that read k and write m arrays of 64 MiB")
    ap.add_argument("--traffic-table", nargs=2, metavar=("READ_DIR", "WRITE_DIR"),
                    help="per-kernel HBM bytes from two rocprofv3 --pmc output folders")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="time a plain 1 GiB copy kernel and exit")
    ap.add_argument("--ensemble", action="store_true",
                    help="the 48-grid ensemble workload instead of the 256^3 grid")
    ap.add_argument("--ens-shrink", type=int, default=1, help="divide ensemble grid sizes by this")
    ap.add_argument("--learned", type=int, nargs="?", const=1000, default=None, metavar="STEPS",
                    help="the learned configuration at 192^3: flexible CG with the learned stage "
                         "against CG + Jacobi over STEPS steps (default 1000)")
    ap.add_argument("--network", action="store_true",
                    help="time the learned part's forward and energy-gradient passes alone (--grid, "
                         "--steps, --warmup)")
    a = ap.parse_args()
    if a.traffic_table:
        traffic_table(*a.traffic_table, grid=a.grid)
        return
    if a.calibrate_streams:
        calibrate_streams()
        return
    if a.network:
        bench_network(a)
        return
    if a.learned:
        bench_learned(a)
        return
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.ensemble:
        if a.gpus != 1:
            print("bench.py: one GPU only", file=sys.stderr)
            sys.exit(2)
        bench_ensemble(a)
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the hot path on one GPU only; multi-GPU runs are "
              "not part of it (see DESIGN.md for how the grid would shard)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from mizuho_edit_amd import io
    from mizuho_edit_amd.wpfaw3d import solver_opts

    n = a.grid
    log(f"building the {n}^3 workload (seed {a.seed})")
    case = io.workload("plain-256", seed=a.seed, grid=(n, n, n))
    ncell = n ** 3
    plain_opts = solver_opts("mgcg", rtol=1e-10)

    log("plain FP64 path")
    g64, wall64, its64, _, _, steps64 = run_path(case, plain_opts, a.warmup, a.steps, a.run_steps)
    p64 = g64.download("pressure")
    g64.close()

    from mizuho_edit_amd import learned
    from mizuho_edit_amd._lib import (FLAG_CHEBYSHEV, FLAG_EXTRAPOLATE, FLAG_F32_STENCIL,
                                      FLAG_LEARNED_WARM, FLAG_OP16, FLAG_PROJECT)
    weights = a.weights or learned.WEIGHTS
    net = None
    if not a.no_learned:
        if not os.path.exists(weights):
            log(f"no network weights at {weights}: run `python -m mizuho_edit_amd.learned train` "
                "or pass --no-learned")
            sys.exit(1)
        net = learned.PCNet()
        net.load(weights)
    log("headline path (learned stage + FP32 MG-PCG in FP64 refinement)" if net else
        "headline path without the network (FP32 MG-PCG in FP64 refinement)")
    flags = FLAG_EXTRAPOLATE | (FLAG_LEARNED_WARM if net else 0)
    if not a.jacobi:
        flags |= FLAG_CHEBYSHEV | FLAG_OP16
    if not a.no_project:
        flags |= FLAG_PROJECT
    if a.f32_stencils:
        flags |= FLAG_F32_STENCIL
    fast_opts = solver_opts("mixed", rtol=1e-10, inner_rtol=1e-4, mg_levels=a.fast_levels, flags=flags)
    gf, wallf, itsf, state0, prof, stepsf = run_path(case, fast_opts, a.warmup, a.steps, a.run_steps,
                                                     a.kernel, net)
    pf = gf.download("pressure")
    stage_steps = gf.stage_steps
    stage_timed = [t for t in stage_steps if a.warmup < t <= a.warmup + a.steps]
    gf.close()
    perr = float(np.abs(pf - p64).max() / np.abs(p64).max())

    sps = a.steps / wallf
    sps64 = a.steps / wall64
    total_ms, count = prof
    avg_s = total_ms / 1e3 / max(count, 1)
    bpc = bytes_per_cell(a.kernel, 4, 4 if a.jacobi else 2, not a.jacobi, count, itsf)
    achieved = bpc * ncell / avg_s / 1e9 if count else 0.0
    run_len = len(stepsf)
    whole = {"steps": run_len,
             "fast_steps_per_s": run_len / sum(stepsf),
             "plain_fp64_steps_per_s": run_len / sum(steps64),
             "speedup_vs_plain_fp64": sum(steps64) / sum(stepsf),
             "fast_windows": window_rates(stepsf),
             "plain_fp64_windows": window_rates(steps64)}

    cpu, vs_oracle = None, None
    if not a.no_cpu:
        nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        log(f"CPU oracle: one step with each solver at {n}^3 on {nthreads} threads")
        recs, ref = cpu_baseline(case, state0, nthreads)
        b, c = recs["bicgstab_sor"], recs["cg_jacobi"]
        cpu = {"value": 1.0 / b["seconds"], "unit": "time-steps/s", "cores": nthreads, "kind": "port",
               "sample": f"step {state0['itime'] + 1} of the {n}^3 workload from the GPU timed window's "
                         f"start state, FP64 C restatement on {nthreads} threads, the reference's solver "
                         f"BiCGStab + SOR(10, omega 1.7; {nthreads} SOR blocks) rtol 1e-10 maxit 1000: "
                         f"{b['iterations']} iterations, {'converged' if b['converged'] else 'NOT converged'}, "
                         f"{b['seconds']:.1f} s",
               "solvers": {"bicgstab_sor": {**b, "steps_per_s": 1.0 / b["seconds"]},
                           "cg_jacobi": {**c, "steps_per_s": 1.0 / c["seconds"]}}}
        # the same step on both GPU paths, against the oracle's converged step
        vs_oracle = {"step": state0["itime"] + 1,
                     "oracle": f"BiCGStab + SOR, rtol 1e-10, {b['iterations']} iterations"}
        one_fast = solver_opts("mixed", rtol=1e-10, inner_rtol=1e-4, mg_levels=a.fast_levels,
                               flags=flags & ~FLAG_LEARNED_WARM)
        for name, o in (("fast", one_fast), ("plain_fp64", plain_opts)):
            _, fl = gpu_one_step(case, state0, o)
            vs_oracle[name] = {f: float(np.abs(fl[f][1:-1, 1:-1, 1:-1] - getattr(ref, f)[1:-1, 1:-1, 1:-1]).max()
                                        / np.abs(getattr(ref, f)[1:-1, 1:-1, 1:-1]).max())
                               for f in ("pressure", "phase", "concent")}

    out = {
        "metric": METRIC,
        "value": sps,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wallf / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32" if a.f32_stencils else "f64") + " stencils (f64 fields and update) + "
                 "f32 solve in f64 refinement",
        "data": "synthetic: seeded Gaussian random field rock (corr 4 cells), Kozeny-Carman k",
        "config": {"workload": f"wormholing {n}^3, steps {a.warmup + 1}..{a.warmup + a.steps} "
                               f"of a {run_len}-step run", "grid": [n, n, n],
                   "solver": ("" if a.no_project else "projected initial guess (last 5 solution changes) + ") +
                             ("learned first stage (gated) + " if net else "") +
                             (f"FP32 MG-PCG ({a.fast_levels} levels, 2 Jacobi sweeps)" if a.jacobi else
                              f"FP32 MG-PCG ({a.fast_levels} levels, degree-2 Chebyshev, FP16-stored fine faces)") +
                             " in FP64 refinement, rtol 1e-10",
                   "plain_solver": "FP64 MG-PCG (5 levels, 2 Jacobi sweeps), rtol 1e-10",
                   "parallelism": "1 GPU"},
        "plain_fp64_steps_per_s": sps64,
        "speedup_vs_plain_fp64": sps / sps64,
        "whole_run": whole,
        "p_rel_err_vs_plain_fp64": perr,
        "rel_err_vs_fp64_oracle_one_step": vs_oracle,
        "cell_updates_per_s": sps * ncell,
        "iterations_per_step": {"plain_fp64": its64 / a.steps, "fast_fp32_inner": itsf / a.steps},
        "learned_stage": ({"weights": os.path.relpath(weights, ROOT),
                           "solves_with_stage": len(stage_steps),
                           "timed_solves_with_stage": len(stage_timed),
                           "steps_with_stage": stage_steps[:40],
                           "gate": "the stage runs in the first and third of every 100 solves (probes, against "
                                   "the second and fourth without it) and otherwise only while the mean solve "
                                   "time with it over the probes is at least 3% below the other's",
                           "contribution": ("none in the timed steps: the stage did not run there"
                                            if not stage_timed else
                                            f"ran in {len(stage_timed)} of the {a.steps} timed solves")}
                          if net else None),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(a.kernel, n) if a.kernel == "cg_update" and not a.jacobi
                     else None,
                     "kernel": a.kernel, "launches": count, "avg_us": avg_s * 1e6,
                     "bytes_per_cell": bpc,
                     "frac_of_streaming_probe": achieved / STREAM_GBS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
