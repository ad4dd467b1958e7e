"""Throughput of the IMEX_SfloW2D per-timestep path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload volcano]

One step = one full timestep of the 2048 x 2048 grid (check_solve mask,
stencil pass with the CFL reduction, dt policy, stage-2 Newton and assembly,
erosion/deposition, post-step diagnostics, topography update) with the state
resident in HBM.  The default workload is the synthetic volcano: 2048^2
cells, dx = 5 m, cone + 32 bumps + fractal noise, radial source 1e6 kg/s,
three solid classes, erosion/deposition with topography change, 1000
timesteps from t = 0.  ``--workload wet`` runs the fully wet 2048^2 incline
variant (every cell active) that exposes the stencil and Newton kernels.

Printed: one JSON line.  ``value`` is grid cell-updates/s (nx*ny*steps /
wall time of the timed steps, run as the library runs them: captured graphs
replayed on its stream); the active variant (sum of solve_cells / time) is
in ``config``.  ``roofline`` prices the dominant kernel's algorithmic HBM
bytes per launch against 8 TB/s using the kernel's average duration,
measured with HIP events on the library's stream in a second pass of the
same steps (per-kernel events need direct launches, which slow the step, so
that pass's wall time is reported separately); ``traffic`` comes from the
committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of that kernel when one
exists.  ``cpu_baseline`` times the C oracle (the reference
algorithm restated, OpenMP over cells, reference interface semantics) on the
host cores over the first steps of the same workload.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per processed cell of each kernel (doubles unless
# stated): what the kernel has to read and write at least once.
#   stencil: q (NV) + B'x, B'y + mask byte in, D (NV) out, plus the slope
#            correction / curvature fields when they are on
#   local:   q, D, B'x, B'y, mask in; q out; erodible/deposit/erosion
#            (read+write, NS each) when erosion or settling is on; B_cent
#            read+write when the topography changes; hmax, pdynmax,
#            mod_vel_max, vulnerability word read+write
#   mask:    state-flag byte in (+ source byte with a radial source), mask
#            byte out (every grid cell)
#   topo:    B_cent in, B'x, B'y out (cells near the flow)


def kernel_bytes(p, name):
    ns = p.n_solid
    nv = 4 + ns
    if name == "stencil":
        b = 8 * nv + 16 + 1 + 8 * nv
        if p.slope_correction_flag:
            b += 8 * 4
        if p.curvature_term_flag:
            b += 8 * 3
        return b
    if name == "local":
        b = 8 * nv * 2 + 16 + 1 + 8 * nv
        if p.erosion_coeff > 0.0 or p.settling_flag:
            b += 3 * ns * 16
        if p.topo_change_flag:
            b += 16
        b += 3 * 16 + 8
        return b
    if name == "mask":
        return 2 + (1 if p.radial_source_flag else 0)
    if name == "topo":
        return 24
    return 0


def step_bytes(p):
    """Bytes per active cell-update of the whole step (stencil + local)."""
    return kernel_bytes(p, "stencil") + kernel_bytes(p, "local")


def parse_traffic(kernel, workload):
    """HBM bytes per dispatch of `kernel` from the committed rocprofv3
    counter summary: FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by
    what the same counter reads for a device copy of known size (the file
    records the raw values, the factors and the product)."""
    path = os.path.join(HERE, "profiles", "%s_%s_traffic.txt" % (kernel, workload))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            m = re.match(r"\s*hbm_bytes_per_dispatch\s*=\s*([0-9.eE+-]+)", line)
            if m:
                return float(m.group(1))
    return None


def make_deck(workload):
    from imex_sflow2d_v2_amd import decks
    if workload == "volcano":
        return decks.volcano_2048()
    if workload == "wet":
        return decks.wet_2048()
    raise SystemExit("unknown workload %r" % workload)


def cpu_baseline(workload, steps, threads):
    """Time the oracle (reference interface semantics) on the host cores
    over the first `steps` timesteps of the workload."""
    from oracle.pyoracle import Oracle
    d = make_deck(workload)
    o = Oracle(d.params, fresh=False, threads=threads).load_deck(d)
    t0 = time.perf_counter()
    n = o.run(steps, dt0=d.dt0)
    el = time.perf_counter() - t0
    c = o.counters()
    o.close()
    cells = d.params.nx * d.params.ny
    return {
        "value": cells * n / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": "first %d timesteps of the %s workload (%dx%d) from t=0 on %d "
                  "host threads, %.2f s wall; active cell-updates/s %.4g"
                  % (n, workload, d.params.nx, d.params.ny, threads, el,
                     c["active_cell_updates"] / el),
    }


def bench_ensemble(args):
    """Batched ensemble: `members` grids of log-uniform sizes 128^2..1024^2
    with sampled parameters, `steps` (default 500) steps each, advanced by
    batched launches in captured graphs after the strip-height autotuner has
    picked the shape on a warm-up copy; compared with a sequential loop that
    advances the members one after the other, each in its own context."""
    from imex_sflow2d_v2_amd import decks
    from imex_sflow2d_v2_amd.solver import Ensemble, Solver
    steps = args.steps if args.steps != 1000 else 500
    ds = decks.ensemble_members(count=args.members)
    params = [d.params for d in ds]
    cells = sum(p.nx * p.ny for p in params)
    dt0 = ds[0].dt0
    warm = Ensemble(params).load(ds)
    best = warm.autotune(max(1, args.warmup // 3 or 1), dt0)
    warm.close()
    ens = Ensemble(params).load(ds)
    ens.set_tuning(best, 25)
    for m in ens.members:
        m.sync()
    t0 = time.perf_counter()
    n = ens.advance(steps, dt0)
    el = time.perf_counter() - t0
    dev_ms = ens.last_ms()
    iters = sum(m.stats()["newton_iters"] for m in ens.members)
    ncell = sum(m.stats()["newton_cells"] for m in ens.members)
    ens.close()
    # sequential reference loop: one context per member, one after the other
    seq = 0.0
    for d in ds:
        g = Solver(d.params).load_deck(d)
        g.sync()
        a = time.perf_counter()
        g.advance(steps, dt0=d.dt0)
        g.sync()
        seq += time.perf_counter() - a
        g.close()
    return {
        "metric": "ensemble members/s (%d steps each)" % steps,
        "value": len(ds) / el,
        "unit": "members/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": el / max(n, 1) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "ensemble_%d" % len(ds),
            "members": len(ds),
            "total_cells": cells,
            "sizes_min_max": [min(p.nx for p in params), max(p.nx for p in params)],
            "strip_rows_autotuned": best,
            "graph_steps": 25,
            "grid_cell_updates_per_s": cells * n / el,
            "device_ms": dev_ms,
            "sequential_members_per_s": len(ds) / seq,
            "speedup_vs_sequential": seq / el,
            "newton_iters_per_cell": iters / max(1, ncell),
        },
        "roofline": None,
        "cpu_baseline": None,
    }


def bench_warmstart(args):
    """Learned Newton warm start: harvest (state entering Newton, converged
    solution) pairs from the solver on the fully wet 2048^2 grid, train the
    MLP on the device, then run the same steps plain and warm-started on the
    fully wet grid and on the volcano from t = 0 and compare Newton
    iterations per cell, fallback rate, step time and the difference of the
    states."""
    from imex_sflow2d_v2_amd import learned
    from imex_sflow2d_v2_amd.solver import Solver
    from imex_sflow2d_v2_amd import decks
    # harvest from the benchmark's configurations: the 1-D dam break, the
    # incline collapse, a volcano on another seeded DEM (the evaluation DEM
    # stays unseen) and one step of the fully wet grid
    sources = [("dambreak_1d", decks.dambreak_1d(), 400),
               ("incline_512", decks.incline_2d(), 100),
               ("volcano_seed5", decks.volcano_2048(seed=5), 1000),
               ("wet_2048", decks.wet_2048(), 1)]
    chunks, produced, t_h, harvest_info = [], 0, 0.0, []
    for name, dd, ns in sources:
        g = Solver(dd.params).load_deck(dd)
        learned.start_harvest(g, 30_000_000)
        t0 = time.perf_counter()
        g.advance(ns, dt0=dd.dt0)
        g.sync()
        pp, nprod = learned.harvested(g)
        t_h += time.perf_counter() - t0
        learned.stop_harvest(g)
        g.close()
        chunks.append(pp)
        produced += nprod
        harvest_info.append({"deck": name, "steps": ns, "pairs": int(pp.shape[0])})
    pairs = np.concatenate(chunks, axis=0)
    del chunks
    m = learned.Mlp(seed=1)
    t0 = time.perf_counter()
    hist = m.train(pairs, epochs=3, batch=8192, lr=1e-3, seed=2)
    t_train = time.perf_counter() - t0
    del pairs
    res = {}
    steps = {"wet": 40, "volcano": 300}
    for wl in ("wet", "volcano"):
        d = make_deck(wl)
        out = {}
        qs = {}
        for mode in ("plain", "warm"):
            s_ = Solver(d.params).load_deck(d)
            if mode == "warm":
                learned.set_warm_start(s_, m)
            s_.advance(2, dt0=d.dt0)   # warm-up (graph capture etc.)
            s_.sync()
            s_.stats(reset=True)
            a = time.perf_counter()
            s_.advance(steps[wl], dt0=d.dt0)
            s_.sync()
            el = time.perf_counter() - a
            st = s_.stats()
            out[mode] = {"ms_per_step": el / steps[wl] * 1e3,
                         "newton_iters_per_cell": st["newton_iters"] / max(1, st["newton_cells"]),
                         "line_search_evals_per_cell": st["line_search_evals"] / max(1, st["newton_cells"]),
                         "warm_cells": st["warm_cells"], "fallbacks": st["warm_fallbacks"],
                         "newton_cells": st["newton_cells"],
                         "nonconverged": st["newton_nonconverged"]}
            qs[mode] = s_.q()
            s_.close()
        rel = 0.0
        for v in range(qs["plain"].shape[0]):
            sc = max(float(np.abs(qs["plain"][v]).max()), 1e-300)
            rel = max(rel, float(np.abs(qs["warm"][v] - qs["plain"][v]).max()) / sc)
        w, p_ = out["warm"], out["plain"]
        res[wl] = {
            "steps": steps[wl],
            "plain": p_, "warm": w,
            "iteration_reduction": p_["newton_iters_per_cell"] / max(w["newton_iters_per_cell"], 1e-30),
            "end_to_end_speedup": p_["ms_per_step"] / w["ms_per_step"],
            "fallback_rate": w["fallbacks"] / max(1, w["newton_cells"]),
            "max_rel_diff_warm_vs_plain": rel,
        }
    return {
        "metric": "Newton iterations per cell, plain / learned warm start (volcano 2048^2)",
        "value": res["volcano"]["iteration_reduction"],
        "unit": "x",
        "n_gpus": 1,
        "steps": steps["volcano"],
        "warmup": 2,
        "ms_per_step": res["volcano"]["warm"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (MLP bf16/fp32)",
        "data": "synthetic",
        "config": {
            "workload": "warmstart",
            "harvest": {"sources": harvest_info, "pairs": int(produced), "seconds": t_h},
            "training": {"epochs": 3, "batch": 8192, "lr": 1e-3, "loss_history": [float(x) for x in hist],
                         "seconds": t_train},
            "results": res,
        },
        "roofline": None,
        "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="volcano",
                    choices=["volcano", "wet", "ensemble", "warmstart"])
    ap.add_argument("--members", type=int, default=256,
                    help="ensemble workload: number of members")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one device-to-device copy of 512 MiB and exit: the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / "
                         "WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the multi-GPU decomposition is a design only)" % args.gpus,
              file=sys.stderr)
        return 2

    from imex_sflow2d_v2_amd.solver import Solver

    if args.copy_probe:
        # torch only here: the measured runs below use the library's own
        # stream and synchronisation
        import torch
        if not torch.cuda.is_available():
            print("bench.py: no GPU visible", file=sys.stderr)
            return 1
        n = 1 << 26  # 2^26 doubles = 512 MiB read + 512 MiB written
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 8 * n, "copy_probe_bytes_written": 8 * n}))
        return 0

    if args.workload == "ensemble":
        print(json.dumps(bench_ensemble(args)), flush=True)
        return 0
    if args.workload == "warmstart":
        print(json.dumps(bench_warmstart(args)), flush=True)
        return 0

    d = make_deck(args.workload)
    p = d.params
    cells = p.nx * p.ny

    # warm-up on its own context (code objects, caches), then a fresh
    # context timed from t = 0
    if args.warmup > 0:
        w = Solver(p).load_deck(d)
        w.advance(args.warmup, dt0=d.dt0)
        w.sync()
        w.close()
    # the timed steps: the library's production path (captured 20-step
    # graphs replayed on its stream)
    g = Solver(p).load_deck(d)
    g.sync()
    g.stats(reset=True)
    g.sync()
    t0 = time.perf_counter()
    n = g.advance(args.steps, dt0=d.dt0)
    g.sync()
    el = time.perf_counter() - t0
    st = g.stats()
    g.close()
    # the same steps again with a HIP event pair around every kernel on the
    # library's stream (direct launches: the events slow the step down, so
    # this pass gives the kernel durations, not the throughput)
    h = Solver(p).load_deck(d)
    h.set_timing(True)
    h.sync()
    t1 = time.perf_counter()
    h.advance(args.steps, dt0=d.dt0)
    h.sync()
    el_timed = time.perf_counter() - t1
    kt = h.kernel_times()
    h.close()

    steps = max(n, 1)
    active = st["cell_updates_active"]
    dom = max(((k, v) for k, v in kt.items() if v[1] > 0), key=lambda kv: kv[1][0])
    dname, (dms, dcalls) = dom
    avg_s = dms / dcalls * 1e-3
    computed = st["cell_updates_computed"]
    units = computed / steps if dname in ("stencil", "local") else cells
    alg_bytes = kernel_bytes(p, dname) * units
    achieved = alg_bytes / avg_s / 1e9
    traffic = parse_traffic(dname, args.workload)

    out = {
        "metric": "FP64 cell-updates/s at 2048^2",
        "value": cells * steps / el,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": el / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "%s_%dx%d" % (args.workload, p.nx, p.ny),
            "nx": p.nx, "ny": p.ny, "n_solid": p.n_solid, "n_rk": p.n_rk,
            "dx": p.cell_size,
            "active_cell_updates_per_s": active / el,
            "active_fraction": active / float(cells * steps),
            "computed_cell_updates_per_s": computed / el,
            "newton_iters_per_cell": st["newton_iters"] / max(1, st["newton_cells"]),
            "newton_iter_hist_1_2_3_4_8_16_32_100": st["newton_iter_hist"],
            "newton_iter_max": st["newton_iter_max"],
            "line_search_evals_per_cell": st["line_search_evals"] / max(1, st["newton_cells"]),
            "simulated_time_s": g.t,
            "kernel_ms_per_step": {k: v[0] / steps for k, v in kt.items() if v[1]},
            "kernel_times_pass_ms_per_step": el_timed / steps * 1e3,
            "step_bytes_per_active_cell": step_bytes(p),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dname,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_unit": kernel_bytes(p, dname),
            "units_per_launch": units,
            "avg_launch_ms": avg_s * 1e3,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_steps, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
