"""Throughput of the IMEX_SfloW2D per-timestep path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload volcano]

One step = one full timestep of the 2048 x 2048 grid (check_solve mask,
stencil pass with the CFL reduction, dt policy, stage-2 Newton and assembly,
erosion/deposition, post-step diagnostics, topography update) with the state
resident in HBM.  The default workload is the synthetic volcano: 2048^2
cells, dx = 5 m, cone + 32 bumps + fractal noise, radial source 1e6 kg/s,
three solid classes, erosion/deposition with topography change, a run of
1000 timesteps from t = 0.  ``--workload wet`` runs the fully wet 2048^2
incline variant (every cell active) that exposes the stencil and Newton
kernels; ``--workload profiles`` a 2048^2 channel of moving flow with the
vertical-profile closures and rheology model 8 (the tableau-driven path).

Timing.  W warm-up steps run on the context that is timed (code objects,
caches, the captured step graphs); the deck is then loaded again and the run
starts from t = 0.  It is timed in segments between synchronisations: the
first K steps, the `develop` steps after them, the timed window of K steps
at a developed state (steps 500..500+K by default), and the rest of the
1000-step run.  ``value`` is grid cell-updates/s of the window (nx*ny*K /
its wall time, the steps run as the library runs them: captured graphs
replayed on its stream); the t = 0 window and the whole run are reported in
``config``, as are active and computed cell-updates/s.

``roofline`` prices the dominant kernel's algorithmic HBM bytes per launch
against 8 TB/s using the kernel's average duration in the same window,
measured with HIP events on the library's stream in a second pass (per-kernel
events need direct launches, which slow the step, so that pass's wall time
is reported separately).  The fully wet 2048^2 grid runs in the same
invocation and its roofline (with rocprofv3 FETCH_SIZE/WRITE_SIZE traffic
from the committed summaries) is in ``config.wet_2048``; the batched
256-member ensemble (members/s) and the 2048^2 vertical-profile channel
(ms per step) are in ``config.ensemble_256`` and ``config.profiles_2048``.
``cpu_baseline``
times the C oracle (the reference algorithm restated, OpenMP over cells,
reference interface semantics) on the host cores, started from the device's
state at the start of the window.
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
    if workload == "profiles":
        # the vertical-profile closures on a 2048^2 channel of moving flow;
        # the initial state converted by the library's own qp_to_qc
        from imex_sflow2d_v2_amd.solver import Solver
        holder = {}

        def convert(p, qp):
            s = holder.setdefault("s", Solver(p))
            return s.qp_to_qc(qp)
        d = decks.profile_channel(nx=2048, ny=2048, dx=5.0, convert=convert)
        holder["s"].close()
        return d
    raise SystemExit("unknown workload %r" % workload)


def host_cpu():
    """Model name and physical core count of the host CPU (/proc/cpuinfo)."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return model, (len(cores) or None)


def cpu_baseline(workload, steps, threads, start=None, device_q=None):
    """Time the oracle (reference interface semantics) on the host cores
    over `steps` timesteps of the workload: from t = 0, or from the state
    `start` (q, erodible, B_cent, t, dt) that the device reached, so that
    the sample covers the same window as the timed device steps.  With
    `device_q` (the device's state after the same steps from t = 0) the
    oracle also checks it: the largest relative difference per conserved
    variable is reported."""
    from oracle.pyoracle import Oracle
    d = make_deck(workload)
    o = Oracle(d.params, fresh=False, threads=threads).load_deck(d)
    dt0, t_start = d.dt0, 0.0
    if start is not None:
        q, erodible, B, t_start, dt0 = start
        o.lib.orc_set_topography(o.h, np.asfortranarray(B).ctypes.data_as(
            o.lib.orc_set_topography.argtypes[1]), None)
        o.set_q(q)
        if d.params.n_solid:
            e = np.asfortranarray(erodible)
            o.lib.orc_set_erodible(o.h, e.ctypes.data_as(o.lib.orc_set_erodible.argtypes[1]))
        o.t = t_start
    t0 = time.perf_counter()
    n = o.run(steps, dt0=dt0)
    el = time.perf_counter() - t0
    c = o.counters()
    diff = None
    if device_q is not None:
        qo = o.q()
        diff = max(float(np.abs(device_q[v] - qo[v]).max() / max(np.abs(qo[v]).max(), 1e-300))
                   for v in range(qo.shape[0]))
    o.close()
    cells = d.params.nx * d.params.ny
    model, phys = host_cpu()
    where = ("from t = %.4g s (the device's state at the start of the timed window)" % t_start
             if start is not None else "from t = 0")
    return {
        "value": cells * n / el,
        "unit": "cell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d timesteps of the %s workload (%dx%d) %s on %d host threads, %.2f s wall; "
                  "active cell-updates/s %.4g"
                  % (n, workload, d.params.nx, d.params.ny, where, threads, el,
                     c["active_cell_updates"] / el),
        "host_cpu": model,
        "host_physical_cores": phys,
        "active_cell_updates_per_s": c["active_cell_updates"] / el,
        "max_rel_diff_device_vs_oracle": diff,
    }


def bench_ensemble(args):
    """Batched ensemble: `members` grids of log-uniform sizes 128^2..1024^2
    with sampled parameters, `steps` (default 500) steps each, advanced by
    batched launches in captured graphs after the strip-height autotuner has
    picked the shape on a warm-up copy; compared with a sequential loop that
    advances the members one after the other, each in its own context."""
    from imex_sflow2d_v2_amd import decks
    from imex_sflow2d_v2_amd.solver import Ensemble, Solver
    steps = args.ens_steps
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


def extra_ensemble(steps, members=256, warmup=6):
    """The batched ensemble (members of 128^2..1024^2, `steps` steps each)
    for the volcano line's config: members/s of the batched launches after
    the strip-height autotuner, without the sequential comparison."""
    from imex_sflow2d_v2_amd import decks
    from imex_sflow2d_v2_amd.solver import Ensemble
    ds = decks.ensemble_members(count=members)
    params = [d.params for d in ds]
    cells = sum(p.nx * p.ny for p in params)
    dt0 = ds[0].dt0
    warm = Ensemble(params).load(ds)
    best = warm.autotune(max(1, warmup // 3), dt0)
    warm.close()
    ens = Ensemble(params).load(ds)
    ens.set_tuning(best, 25)
    ens.advance(25, dt0)
    ens.load(ds)
    for m in ens.members:
        m.sync()
    t0 = time.perf_counter()
    n = ens.advance(steps, dt0)
    for m in ens.members:
        m.sync()
    el = time.perf_counter() - t0
    ens.close()
    return {"members": len(ds), "steps_each": n, "members_per_s": len(ds) / el,
            "member_steps_per_s": len(ds) * n / el,
            "grid_cell_updates_per_s": cells * n / el, "total_cells": cells,
            "strip_rows_autotuned": best}


def extra_profiles(steps):
    """The vertical-profile channel at 2048^2 (rheology model 8, two
    classes) for the volcano line's config: ms per step through the step
    graphs after a warm-up that captures them."""
    from imex_sflow2d_v2_amd.solver import Solver
    d = make_deck("profiles")
    p = d.params
    h = Solver(p).load_deck(d)
    h.advance(20, dt0=d.dt0)
    h.load_deck(d)
    h.sync()
    t0 = time.perf_counter()
    n = h.advance(steps, dt0=d.dt0)
    h.sync()
    el = time.perf_counter() - t0
    h.close()
    return {"steps": n, "ms_per_step": el / max(n, 1) * 1e3,
            "cell_updates_per_s": p.nx * p.ny * n / el}


def bench_warmstart(args):
    """Newton starting points on the 2048^2 volcano (and the fully wet grid):
    the plain predictor q_si (the reference's), the learned warm start (an
    MLP trained on the device on pairs the solver harvests from the
    configurations: the 1-D dam break, the 512^2 incline, a volcano on
    another seeded DEM and the fully wet grid), and the closed-form Voellmy
    start.  Newton iterates to the reference tolerance from each; reported:
    iterations per cell, fallback rate, step time through the captured
    graphs and the largest difference to the plain run's state."""
    from imex_sflow2d_v2_amd import learned
    from imex_sflow2d_v2_amd.solver import Solver
    from imex_sflow2d_v2_amd import decks
    sources = [("dambreak_1d", decks.dambreak_1d(), 2300),
               ("incline_512", decks.incline_2d(), args.harvest_incline_steps),
               ("volcano_seed5", decks.volcano_2048(seed=5), 1000),
               ("wet_2048", decks.wet_2048(), 2)]
    chunks, produced, t_h, harvest_info = [], 0, 0.0, []
    for name, dd, ns in sources:
        g = Solver(dd.params).load_deck(dd)
        learned.start_harvest(g, 45_000_000)
        t0 = time.perf_counter()
        g.advance(ns, dt0=dd.dt0, t_end=dd.t_end if dd.t_end else 1e30)
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
    hist = m.train(pairs, epochs=args.train_epochs, batch=8192, lr=1e-3, seed=2)
    t_train = time.perf_counter() - t0
    n_pairs = int(pairs.shape[0])
    del pairs
    res = {}
    steps = {"volcano": 300, "wet": 20}
    for wl in ("volcano", "wet"):
        d = make_deck(wl)
        out, qs = {}, {}
        for mode in ("plain", "learned", "closed_form"):
            s_ = Solver(d.params).load_deck(d)
            if mode == "learned":
                learned.set_warm_start(s_, m)
            elif mode == "closed_form":
                s_.set_closed_form_start(True)
            s_.advance(2 * 20, dt0=d.dt0)   # capture the graphs
            s_.load_deck(d)
            s_.sync()
            s_.stats(reset=True)
            a = time.perf_counter()
            s_.advance(steps[wl], dt0=d.dt0)
            s_.sync()
            el = time.perf_counter() - a
            st = s_.stats()
            out[mode] = {"ms_per_step": el / steps[wl] * 1e3,
                         "newton_iters_per_cell": st["newton_iters"] / max(1, st["newton_cells"]),
                         "newton_iter_hist_1_2_3_4_8_16_32_100": st["newton_iter_hist"],
                         "line_search_evals_per_cell": st["line_search_evals"] / max(1, st["newton_cells"]),
                         "started": st["warm_cells"], "fallbacks": st["warm_fallbacks"],
                         "newton_cells": st["newton_cells"],
                         "fallback_rate": (st["warm_fallbacks"] / max(1, st["newton_cells"])
                                           if mode != "plain" else None),
                         "nonconverged": st["newton_nonconverged"],
                         "sparse_steps": st["sparse_steps"]}
            qs[mode] = s_.q()
            s_.close()
        p_ = out["plain"]
        for mode in ("learned", "closed_form"):
            rel = 0.0
            for v in range(qs["plain"].shape[0]):
                sc = max(float(np.abs(qs["plain"][v]).max()), 1e-300)
                rel = max(rel, float(np.abs(qs[mode][v] - qs["plain"][v]).max()) / sc)
            w = out[mode]
            w["iteration_reduction"] = p_["newton_iters_per_cell"] / max(w["newton_iters_per_cell"], 1e-30)
            w["end_to_end_speedup"] = p_["ms_per_step"] / w["ms_per_step"]
            w["max_rel_diff_vs_plain"] = rel
        res[wl] = {"steps": steps[wl], **out}
    lw = res["volcano"]["learned"]
    return {
        "metric": "Newton iterations per cell, plain / learned warm start (volcano 2048^2)",
        "value": lw["iteration_reduction"],
        "unit": "x",
        "n_gpus": 1,
        "steps": steps["volcano"],
        "warmup": 40,
        "ms_per_step": lw["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (MLP bf16/fp32)",
        "data": "synthetic",
        "config": {
            "workload": "warmstart",
            "harvest": {"sources": harvest_info, "pairs": n_pairs, "produced": int(produced),
                        "seconds": t_h},
            "training": {"epochs": args.train_epochs, "batch": 8192, "lr": 1e-3,
                         "loss_history": [float(x) for x in hist], "seconds": t_train},
            "results": res,
        },
        "roofline": None,
        "cpu_baseline": None,
    }


def roofline(p, kt, st, steps, traffic_workload=None):
    """Roofline object of the dominant kernel of a timing pass: algorithmic
    bytes per launch (bytes per processed cell x cells processed per launch)
    over the kernel's mean launch duration from the HIP events."""
    cells = p.nx * p.ny
    dname, (dms, dcalls) = max(((k, v) for k, v in kt.items() if v[1] > 0),
                               key=lambda kv: kv[1][0])
    avg_s = dms / dcalls * 1e-3
    units = st["cell_updates_computed"] / steps if dname in ("stencil", "local") else cells
    achieved = kernel_bytes(p, dname) * units / avg_s / 1e9
    return {
        "bound": "hbm",
        "kernel": dname,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": parse_traffic(dname, traffic_workload) if traffic_workload else None,
        "bytes_per_unit": kernel_bytes(p, dname),
        "units_per_launch": units,
        "avg_launch_ms": avg_s * 1e3,
    }


def bench_wet_roofline(nsteps):
    """The fully wet 2048^2 grid (every cell active and computed): the
    HBM-roofline case of the plain path.  Steps per second through the
    captured graphs, and each big kernel's achieved bytes over its mean
    duration from a timing pass."""
    from imex_sflow2d_v2_amd.solver import Solver
    d = make_deck("wet")
    p = d.params
    g = Solver(p).load_deck(d)
    g.advance(2 * 20, dt0=d.dt0)  # capture the graphs
    g.load_deck(d)
    g.sync()
    g.stats(reset=True)
    a = time.perf_counter()
    n = g.advance(nsteps, dt0=d.dt0)
    g.sync()
    el = time.perf_counter() - a
    st = g.stats()
    g.close()
    h = Solver(p).load_deck(d)
    h.set_timing(True)
    h.advance(nsteps, dt0=d.dt0)
    h.sync()
    kt = h.kernel_times()
    h.close()
    kernels = {}
    for name in ("local", "stencil", "mask"):
        if name in kt and kt[name][1]:
            ms = kt[name][0] / kt[name][1]
            units = st["cell_updates_computed"] / max(n, 1) if name != "mask" else p.nx * p.ny
            ach = kernel_bytes(p, name) * units / (ms * 1e-3) / 1e9
            kernels[name] = {"avg_launch_ms": ms, "achieved_GBs": ach,
                             "frac": ach / HBM_PEAK_GBS, "bytes_per_unit": kernel_bytes(p, name),
                             "traffic": parse_traffic(name, "wet")}
    return {
        "steps": n,
        "ms_per_step": el / max(n, 1) * 1e3,
        "grid_cell_updates_per_s": p.nx * p.ny * n / el,
        "newton_iters_per_cell": st["newton_iters"] / max(1, st["newton_cells"]),
        "roofline": roofline(p, kt, st, max(n, 1), traffic_workload="wet"),
        "kernels": kernels,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="timed steps (the window at a developed state)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--develop", type=int, default=None,
                    help="untimed steps from t = 0 before the timed window "
                         "(default 500 for the volcano, 0 for the wet grid)")
    ap.add_argument("--total", type=int, default=1000,
                    help="length of the whole run the window is part of (the "
                         "volcano configuration's 1000 timesteps); 0: stop after "
                         "the window")
    ap.add_argument("--wet-steps", type=int, default=10,
                    help="steps of the fully wet 2048^2 grid run for its roofline")
    ap.add_argument("--no-wet", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="volcano workload: skip the ensemble and profile-path lines of config")
    ap.add_argument("--extra-ens-steps", type=int, default=100,
                    help="steps per member of the ensemble line in the volcano config")
    ap.add_argument("--workload", default="volcano",
                    choices=["volcano", "wet", "profiles", "ensemble", "warmstart"])
    ap.add_argument("--ens-steps", type=int, default=500,
                    help="ensemble workload: steps per member")
    ap.add_argument("--harvest-incline-steps", type=int, default=150,
                    help="warmstart workload: harvested steps of the 512^2 incline")
    ap.add_argument("--train-epochs", type=int, default=3,
                    help="warmstart workload: training epochs")
    ap.add_argument("--members", type=int, default=256,
                    help="ensemble workload: number of members")
    ap.add_argument("--cpu-steps", type=int, default=100)
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
    develop = args.develop if args.develop is not None else (500 if args.workload == "volcano" else 0)
    K = args.steps

    # Warm-up on the context that is then timed: code objects, caches and the
    # captured step graphs (the develop steps below capture them too).
    # Loading the deck again starts a new run from t = 0 with the same
    # device pointers, so the graphs stay valid.
    g = Solver(p).load_deck(d)
    if args.warmup > 0:
        g.advance(max(args.warmup, 2 * 20), dt0=d.dt0)
        g.sync()
        g.load_deck(d)
    g.sync()
    # One run from t = 0 in segments, each timed between synchronisations:
    # the first K steps, the develop steps, the timed window of K steps at a
    # developed state (the headline), and the rest up to `total` steps.
    seg = {}
    stats = {}

    def segment(name, n):
        if n <= 0:
            return 0
        g.stats(reset=True)
        g.sync()
        a = time.perf_counter()
        done = g.advance(n, dt0=d.dt0)
        g.sync()
        seg[name] = (done, time.perf_counter() - a)
        stats[name] = g.stats()
        return done

    n0 = segment("t0_window", min(K, develop) if develop else 0)
    segment("develop", develop - n0)
    t_win = g.t
    dt_win = g.get_dt()
    start_state = None
    if not args.no_cpu:
        er = g.erosion_state()
        start_state = (g.q(), er[0], er[3], t_win, dt_win)
    n = segment("window", K)
    rest = args.total - develop - K if args.total else 0
    segment("rest", rest)
    t_final = g.t
    g.close()
    el = seg["window"][1]
    st = stats["window"]

    # The timed window again with a HIP event pair around every kernel on the
    # library's stream (direct launches: the events slow the step down, so
    # this pass gives the kernel durations, not the throughput).
    h = Solver(p).load_deck(d)
    h.set_timing(True)
    if develop:
        h.advance(develop, dt0=d.dt0)
    h.sync()
    h.set_timing(True)  # clears the develop steps' kernel times
    t1 = time.perf_counter()
    h.advance(K, dt0=d.dt0)
    h.sync()
    el_timed = time.perf_counter() - t1
    kt = h.kernel_times()
    h.close()

    steps = max(n, 1)
    # traffic: the committed counter summary of the same stretch (the volcano
    # window at steps 500-600, or a whole run of a workload without one)
    if develop:
        tw = "volcano_window" if (args.workload == "volcano" and develop == 500) else None
    else:
        tw = args.workload
    rl = roofline(p, kt, st, steps, traffic_workload=tw)
    active = st["cell_updates_active"]
    computed = st["cell_updates_computed"]
    cfg = {
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
        "kernel_ms_per_step": {k: v[0] / steps for k, v in kt.items() if v[1]},
        "kernel_times_pass_ms_per_step": el_timed / steps * 1e3,
        "step_bytes_per_active_cell": step_bytes(p),
        "simulated_time_s": t_final,
    }
    for name in ("t0_window", "rest"):
        if name in seg:
            nn, tt = seg[name]
            cfg[name + "_ms_per_step"] = tt / max(nn, 1) * 1e3
    tot_n = sum(v[0] for v in seg.values())
    tot_t = sum(v[1] for v in seg.values())
    cfg["whole_run"] = {"steps": tot_n, "ms_per_step": tot_t / max(tot_n, 1) * 1e3,
                        "grid_cell_updates_per_s": cells * tot_n / tot_t}
    cfg["window"] = ("steps %d-%d of one run from t = 0 (t = %.4g s at the window's start); "
                     "step graphs captured before the window"
                     % (develop, develop + steps, t_win))
    if args.workload == "volcano" and not args.no_wet:
        cfg["wet_2048"] = bench_wet_roofline(args.wet_steps)
    if args.workload == "volcano" and not args.no_extra:
        # the next workloads of the same path, measured in the same invocation
        cfg["ensemble_256"] = extra_ensemble(args.extra_ens_steps)
        cfg["profiles_2048"] = extra_profiles(20)
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
        "config": cfg,
        "roofline": rl,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        # the profile closures make the oracle's cell-updates about 30x
        # dearer than the plain path's: its sample is --cpu-steps steps
        csteps = args.cpu_steps if args.workload == "profiles" else max(args.cpu_steps, K)
        dq = None
        if args.workload == "profiles":
            # the device's state after the sample's steps, for the oracle to check
            v = Solver(p).load_deck(d)
            v.advance(csteps, dt0=d.dt0)
            dq = v.q()
            v.close()
        out["cpu_baseline"] = cpu_baseline(args.workload, csteps, threads, device_q=dq,
                                           start=start_state if develop else None)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
