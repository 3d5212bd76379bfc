"""FDTD throughput benchmark on one MI355X: Mcell-updates/s and % of the HBM roofline, the
batched sweep's speedup over sequential runs, and the adjoint's backward/forward time.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload layered|boxes|plain|sweep|adjoint]

A step is one FDTD time step (H and E curl updates with CPML, sparse fixes, port/probe
observers) over the whole grid. The default workload is the single large throughput run:
768x768x384 nodes (226.5 M), 1 mm, 16-cell CPML, eps_r ~ U[1,4] per 8-cell z-layer, two
50-ohm z-ports. Inputs are resident in HBM before the timed region starts. The default run
prints one JSON line with:
  * value, ms_per_step: K timed steps of the layered run;
  * roofline: the dominant kernel (the two-step march), HIP-event timed on its stream, against
    8 TB/s, with its measured HBM traffic per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE);
  * sweep: the batched sweep (96 seeded runs of uneven grids, 48-200 nodes per axis, CPML 8,
    1-4 ports, 800-3000 steps each), first one after another through the single-run engine,
    then together through fg_run_batch (autotuning included): both wall times, the speedup,
    whether every record is bit-identical, and the batch step's roofline;
  * adjoint: the inverse-design gradient (192x192x112 nodes, 0.5 mm, CPML 12, a 64x64x16-cell
    design box with eps_r ~ U[1,4], J = |DFT of Ez at a far probe at 8 GHz|^2 over 3000 steps,
    snapshots every 100 steps): forward and backward times, their ratio, the adjoint step's
    roofline and the gradient against central differences (h = 1e-3) of the f64 forward on
    --fd-cells seeded cells;
  * cpu_baseline: the oracle's FP32 restatement of the reference CPU path on the host cores,
    on the full 768x768x384 grid (20 steps after 2 warm-up steps).
--workload sweep and --workload adjoint print those legs as lines of their own (with their own
CPU baseline); --legs none leaves them out of the default line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-updates/s and % HBM roofline at 1 GPU; batched sweep speedup; adjoint/forward time"
HBM_PEAK_GBS = 8000.0


def build_workload(name):
    from fieldgrid_amd import workloads as W
    if name == "layered":
        sim = W.layered()
        pulse = W.layered_pulse()
        extras_kw = {}
        config = {"workload": "layered throughput run", "dims": [768, 768, 384], "dx_mm": 1.0, "pml": 16,
                  "layers": "eps_r~U[1,4] per 8-cell z-layer, seed 3", "ports": 2}
    elif name == "boxes":
        sim = W.boxes()
        pulse = W.boxes_pulse()
        extras_kw = {"boxes": True}
        config = {"workload": "boxes + Debye + port + NTFF", "dims": [256, 256, 256], "dx_mm": 0.5, "pml": 12,
                  "boxes": 40, "debye_boxes": 10, "ntff_freqs_ghz": [2, 5, 10]}
    elif name == "plain":
        sim, probes = W.plain()
        pulse = W.plain_pulse()
        extras_kw = {"probes": probes}
        config = {"workload": "plain vacuum", "dims": [64, 64, 64], "dx_mm": 1.0, "pml": 10, "probes": 16}
    else:
        raise SystemExit(f"unknown workload {name}")
    return sim, pulse, extras_kw, config


def make_extras(sim, steps, extras_kw):
    import math
    from fieldgrid_amd import fdtd as F
    from fieldgrid_amd import workloads as W
    if extras_kw.get("boxes"):
        return W.boxes_extras(sim, steps=steps, check_every=0)
    return F.Extras(max_steps=steps, decay_db=math.inf, probes=extras_kw.get("probes", []))


def cpu_baseline(name, budget_s=20.0):
    """The oracle's FP32 restatement of run_cpu (cpu.rs:424-482) on the host cores. The layered
    run is timed on its full 768x768x384 grid: 20 steps after 2 warm-up steps, the stepping loop
    alone (the oracle's f64 build of the problem is reported beside it). The other workloads
    are timed on a bounded sample of the same recipe. Returns the dict for the JSON line."""
    import math
    from fieldgrid_amd import fdtd as F
    from fieldgrid_amd import workloads as W
    from oracle import oracle as O
    if name == "layered":
        sim = W.layered()
        n = sim.dims()
        nn = n[0] * n[1] * n[2]
        warm, timed = 2, 20
        r = O.run_cpu(sim.desc, 0, W.layered_pulse(), F.Extras(max_steps=warm + timed, decay_db=math.inf), 32,
                      fields=False, warm=warm)
        secs = r.step_seconds
        return {"value": round(nn * timed / secs / 1e6, 3), "unit": "Mcell-updates/s", "cores": O.threads(),
                "kind": "port",
                "sample": f"the full 768x768x384 layered grid: {timed} steps after {warm} warm-up steps in {secs:.1f} s "
                          f"({secs / timed * 1e3:.0f} ms per step; the oracle's build of the problem took "
                          f"{r.build_seconds:.0f} s and is not timed)"}
    if name == "boxes":
        sim = W.boxes(nodes=96)
        pulse = W.boxes_pulse()
        sample = "96^3 version of the boxes recipe (CPML 12, NTFF, Debye)"
        extras = lambda k: W.boxes_extras(sim, steps=k)
    elif name == "adjoint":
        args, _ = W.design(dims=(96, 96, 112), box=((32, 64), (32, 64), (48, 64)), port=(15, 48), probe=(81, 48))
        sim = F.Sim.try_new(args["grid"], args["mats"], None, [], args["ports"], [], args["media"])
        pulse = args["pulse"]
        sample = "96x96x112 version of the design recipe (CPML 12, port, design box)"
        extras = lambda k: F.Extras(max_steps=k, decay_db=math.inf, probes=args["obs"])
    elif name == "sweep":
        r0 = W.sweep_runs(count=1)[0]
        sim, pulse = r0.sim, r0.pulse
        sample = f"the first run of the sweep ({'x'.join(map(str, sim.dims()))})"
        extras = lambda k: F.Extras(max_steps=k, decay_db=math.inf)
    else:
        sim, probes = W.plain()
        pulse = W.plain_pulse()
        sample = "the full 64^3 plain workload"
        extras = lambda k: F.Extras(max_steps=k, decay_db=math.inf, probes=probes)
    n = sim.dims()
    nn = n[0] * n[1] * n[2]
    probe = O.run_cpu(sim.desc, 0, pulse, extras(2), 32, fields=False)  # warm-up and size probe
    per_step = max(probe.step_seconds / 2, 1e-6)
    k = max(2, min(2000, int(budget_s / per_step)))
    r = O.run_cpu(sim.desc, 0, pulse, extras(k), 32, fields=False)
    dt = r.step_seconds  # the stepping loop alone, as on the GPU side
    return {"value": round(nn * r.steps / dt / 1e6, 3), "unit": "Mcell-updates/s", "cores": O.threads(),
            "kind": "port", "sample": f"{sample}: {r.steps} steps of {nn} nodes in {dt:.1f} s"}


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def sweep_leg(args):
    """The batched sweep: sequential single runs, then one fg_run_batch of the same runs.
    Returns (summary for the JSON line, whole-job Mcell-updates/s of the batch, records identical)."""
    import numpy as np
    from fieldgrid_amd import fdtd as F
    from fieldgrid_amd import workloads as W
    t_build = time.perf_counter()
    runs = W.sweep_runs()
    build_s = time.perf_counter() - t_build
    node_steps = sum(int(np.prod(r.sim.dims())) * r.extras.max_steps for r in runs)
    for r in runs:
        r.sim.problem()  # pack every problem before either timed region
    gpu = F.Gpu(0)
    # warm-up: one small run and one small batch load the kernels
    F.run(gpu, runs[0].sim, runs[0].driven, runs[0].pulse, F.Extras(max_steps=4, decay_db=math.inf))
    F.run_batch(gpu, runs[:2], tune=F.Tune(tj=14, seg=16, chunk=1))
    gpu.synchronize()
    log(f"sweep: {len(runs)} runs built in {build_s:.1f} s; sequential runs")
    t0 = time.perf_counter()
    seq = [F.run(gpu, r.sim, r.driven, r.pulse, r.extras) for r in runs]
    gpu.synchronize()
    seq_s = time.perf_counter() - t0
    log(f"sweep: sequential {seq_s:.2f} s; batched run")
    tune = F.Tune(chunk=args.graph_chunk, launch=args.launch)
    t0 = time.perf_counter()
    recs, spectra = F.run_batch(gpu, runs, tune=tune)
    gpu.synchronize()
    bat_s = time.perf_counter() - t0
    log(f"sweep: batched {bat_s:.2f} s")
    same = all(np.array_equal(a.series, b.series) for a, b in zip(seq, recs))
    # algorithmic bytes of one whole-batch step (all runs going) over the tuner's HIP-event time
    # of one step; at two steps per pass this is the effective per-step figure of the pass
    alg = sum(W.bytes_per_node_step(r.sim, None, core=True) * int(np.prod(r.sim.dims())) for r in runs)
    achieved = alg / (tune.step_ms * 1e-3) / 1e9 if tune.step_ms > 0 else None
    depth = tune.used[3]
    if depth == 2 and tune.used_launch == 2:
        kernel = ("k_march2_persist (persistent launches of many two-step passes: each run goes on to its next "
                  "pass when its previous one is done; the pass's last block runs the fixes and observers)")
    elif depth == 2:
        kernel = "k_march2_batch<-1> (class-mixed lists, two steps per pass) + k_fix_batch + k_observe_pair_batch"
    else:
        kernel = "k_march_batch (one launch per class, one step per pass) + k_fix_batch + k_observe_batch"
    # measured HBM bytes per step of a full-width two-step pass, from the committed rocprofv3 PMC
    # summary of this sweep (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_march_sweep_traffic.json")
    if depth == 2 and tune.used_launch != 2 and os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("traffic_bytes_per_step")
    obj = {"runs": len(runs), "node_updates": node_steps, "batched_s": round(bat_s, 3), "sequential_s": round(seq_s, 3),
           "speedup": round(seq_s / bat_s, 3), "batched_mcells": round(node_steps / bat_s / 1e6, 2),
           "sequential_mcells": round(node_steps / seq_s / 1e6, 2), "records_identical": same,
           "build_s": round(build_s, 2),
           "tune": {"tj": tune.used[0], "seg": tune.used[1], "chunk": tune.used[2], "depth": depth,
                    "launch": "persistent" if tune.used_launch == 2 else "per pass",
                    "tune_ms": round(tune.tune_ms, 1), "kept_steps": tune.kept_steps, "from_cache": tune.from_cache},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                        "traffic": traffic, "alg_bytes_per_step": round(alg), "kernel": kernel,
                        "step_ms": round(tune.step_ms, 4),
                        "note": "effective bandwidth: one-step algorithmic bytes of all 96 runs over the "
                                "tuner's HIP-event time per step of the whole batch"}}
    if traffic and tune.step_ms > 0:
        roof = obj["roofline"]
        roof["traffic_gbs"] = round(traffic / (tune.step_ms * 1e-3) / 1e9, 1)
        roof["traffic_frac"] = round(roof["traffic_gbs"] / HBM_PEAK_GBS, 4)
        roof["traffic_over_alg"] = round(traffic / alg, 3)
    return obj, node_steps / bat_s / 1e6, same


def adjoint_leg(args, fd_cells):
    """Forward + discrete adjoint of the inverse-design problem, with a central-difference check
    on `fd_cells` seeded design cells. Returns (summary for the JSON line, forward Mcell/s)."""
    import numpy as np
    from fieldgrid_amd import fdtd as F
    from fieldgrid_amd import workloads as W
    t_build = time.perf_counter()
    a, freq = W.design()
    gpu = F.Gpu(0)
    d = F.Design(gpu, a["grid"], a["mats"], a["pec"], a["lumped"], a["ports"], a["resistive"], a["media"], a["lo"],
                 a["hi"], a["obs"], a["driven"], a["pulse"], a["steps"], 100)
    build_s = time.perf_counter() - t_build
    n = a["grid"].dims()
    nn = n[0] * n[1] * n[2]
    dt = F.time_step(a["grid"])
    d.forward()  # warm-up (kernels loaded, snapshot memory allocated)
    series = d.forward()
    J, dJ = W.dft_objective(series[:, 0], dt, freq)
    grad = d.backward(dJ)
    d.backward(dJ)  # a second, warm backward pass is the one timed
    (fwd_ms, bwd_ms, rec_ms, adj_ms), snaps = d.timing()
    log(f"adjoint: forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms; central differences on {fd_cells} cells")
    fd = None
    if fd_cells:
        box = np.array(a["mats"].eps.reshape(tuple(q - 1 for q in n))[tuple(slice(l, h) for l, h in zip(a["lo"], a["hi"]))],
                       np.float64)
        rng = np.random.default_rng(5)
        flat = rng.choice(box.size, size=fd_cells, replace=False)
        h = 1e-3
        rows = []
        t0 = time.perf_counter()
        for q in flat:
            c = np.unravel_index(int(q), box.shape)
            up, dn = box.copy(), box.copy()
            up[c] += h
            dn[c] -= h
            jp = W.dft_objective(d.forward64(up)[:, 0], dt, freq)[0]
            jm = W.dft_objective(d.forward64(dn)[:, 0], dt, freq)[0]
            rows.append((float(grad[c]), (jp - jm) / (2 * h)))
        g_, f_ = np.array(rows).T
        fd = {"cells": fd_cells, "h": h, "max_rel_err": float(np.max(np.abs(g_ - f_) / np.abs(f_))),
              "max_err_over_max_fd": float(np.max(np.abs(g_ - f_)) / np.max(np.abs(f_))),
              "median_rel_err": float(np.median(np.abs(g_ - f_) / np.abs(f_))),
              "seconds": round(time.perf_counter() - t0, 1)}
    d.close()
    # the same gradient with the whole tape: the forward tapes every step, the backward recomputes
    # nothing (8 B per design edge per step in place of the snapshots)
    w = F.Design(gpu, a["grid"], a["mats"], a["pec"], a["lumped"], a["ports"], a["resistive"], a["media"], a["lo"],
                 a["hi"], a["obs"], a["driven"], a["pulse"], a["steps"], 100, True)
    w.forward()
    w_series = w.forward()
    w.backward(dJ)
    w_grad = w.backward(dJ)
    (w_fwd, w_bwd, _, w_adj), _ = w.timing()
    w.close()
    log(f"adjoint, whole tape: forward {w_fwd:.1f} ms, backward {w_bwd:.1f} ms")
    # the adjoint step's algorithmic bytes: lamE, lamH read and written (48 B/node), the u32
    # coefficient index (every design cell is distinct: 12 B/node), 8 B per psi element, and per
    # design edge its map entry, tape (E^n, curl^n) and f64 dJ/dca, dJ/dcb read-modify-write (44 B)
    sim = F.Sim.try_new(a["grid"], a["mats"], None, [], a["ports"], [], a["media"])
    core = W.bytes_per_node_step(sim, None, core=True)
    box = [h - l for l, h in zip(a["lo"], a["hi"])]
    edges = sum(box[c] * (box[(c + 1) % 3] + 1) * (box[(c + 2) % 3] + 1) for c in range(3))
    alg = core * nn + 44.0 * edges
    adj_step_ms = adj_ms / a["steps"]
    achieved = alg / (adj_step_ms * 1e-3) / 1e9
    # measured HBM bytes per adjoint step, from the committed rocprofv3 PMC summary of this problem
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "k_adj_design_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("traffic_bytes_per_step")
    obj = {"dims": list(n), "steps": a["steps"], "design_box_cells": box, "objective": "|DFT Ez at 8 GHz|^2",
           "snapshot_every": 100, "snapshots": snaps, "forward_ms": round(fwd_ms, 1), "backward_ms": round(bwd_ms, 1),
           "backward_over_forward": round(bwd_ms / fwd_ms, 3), "recompute_ms": round(rec_ms, 1),
           "adjoint_steps_ms": round(adj_ms, 1), "J": J, "grad_max": float(np.max(np.abs(grad))),
           "finite_differences": fd, "build_s": round(build_s, 2),
           "whole_tape": {"forward_ms": round(w_fwd, 1), "backward_ms": round(w_bwd, 1),
                          "backward_over_forward": round(w_bwd / w_fwd, 3), "total_ms": round(w_fwd + w_bwd, 1),
                          "checkpointed_total_ms": round(fwd_ms + bwd_ms, 1),
                          "tape_gb": round(8.0 * edges * a["steps"] / 1e9, 2),
                          "series_identical": bool(np.array_equal(w_series, series)),
                          "grad_identical": bool(np.array_equal(w_grad, grad)),
                          "adjoint_step_gbs": round(alg / (w_bwd / a["steps"] * 1e-3) / 1e9, 1),
                          "adjoint_step_frac": round(alg / (w_bwd / a["steps"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "note": "the forward tapes the design edges of every step; the backward runs the adjoint "
                                  "alone, with no snapshots or recompute (adjoint_step_*: algorithmic bytes of an "
                                  "adjoint step over the backward's time per step)"},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "alg_bytes_per_step": round(alg),
                        "kernel": ("k_adj_march2 (two adjoint steps per launch, per CPML class; HIP-event busy time of "
                                   "its stream per step, the recompute overlapped on a second stream)"),
                        "step_ms": round(adj_step_ms, 4)}}
    if traffic:
        roof = obj["roofline"]
        roof["traffic_gbs"] = round(traffic / (adj_step_ms * 1e-3) / 1e9, 1)
        roof["traffic_frac"] = round(roof["traffic_gbs"] / HBM_PEAK_GBS, 4)
        roof["traffic_over_alg"] = round(traffic / alg, 3)
    return obj, nn * a["steps"] / (fwd_ms * 1e-3) / 1e6


def sweep(args):
    from fieldgrid_amd import fdtd as F
    obj, value, same = sweep_leg(args)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": None, "warmup": 0, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (96 seeded runs)",
        "config": {"workload": "batched sweep", "runs": obj["runs"], "dims": "U[48,200] per axis", "pml": 8,
                   "ports": "1-4, one driven", "steps": "U[800,3000]", "spectra_freqs": 64,
                   "node_updates": obj["node_updates"], "gpu": F.Gpu(0).name},
        "sweep": {k: v for k, v in obj.items() if k != "roofline"},
        "roofline": obj["roofline"],
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline("sweep")
    print(json.dumps(line), flush=True)
    return 0 if same else 1


def adjoint(args):
    from fieldgrid_amd import fdtd as F
    obj, value = adjoint_leg(args, args.fd_cells)
    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": obj["steps"], "warmup": 1, "ms_per_step": round(obj["forward_ms"] / obj["steps"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded design box)",
        "config": {"workload": "adjoint inverse-design gradient", "dims": obj["dims"], "dx_mm": 0.5, "pml": 12,
                   "design_box_cells": obj["design_box_cells"], "objective": obj["objective"],
                   "snapshot_every": 100, "gpu": F.Gpu(0).name},
        "adjoint": {k: v for k, v in obj.items() if k != "roofline"},
        "roofline": obj["roofline"],
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline("adjoint")
    print(json.dumps(line), flush=True)
    return 0


def single(args):
    """One workload through a session: K timed steps, then per-kernel HIP-event timing."""
    from fieldgrid_amd import abi
    from fieldgrid_amd import fdtd as F
    from fieldgrid_amd import workloads as W

    t_build = time.perf_counter()
    sim, pulse, extras_kw, config = build_workload(args.workload)
    build_s = time.perf_counter() - t_build
    n = sim.dims()
    nn = n[0] * n[1] * n[2]
    total = args.warmup + args.steps + args.timed_steps
    extras = make_extras(sim, total, extras_kw)
    extras.engine = abi.FG_ENGINE_SPLIT if args.engine == "split" else abi.FG_ENGINE_FUSED
    bpn = W.bytes_per_node_step(sim, extras)
    gpu = F.Gpu(0)
    log(f"{args.workload}: {nn} nodes built in {build_s:.1f} s; stepping")
    sess = F.Session(gpu, sim, 0, pulse, extras)
    sess.step(args.warmup)
    sess.sync()
    gpu.synchronize()
    t0 = time.perf_counter()
    sess.step(args.steps)
    sess.sync()
    gpu.synchronize()
    secs = time.perf_counter() - t0
    # per-kernel HIP-event timing on the session's stream, after the timed region
    sess.timing(True)
    sess.step(args.timed_steps)
    ms, timed = sess.timing(False)
    energy = sess.energy()
    two = sess.steps_per_pass() == 2  # temporal blocking (march2) for this problem
    sess.close()
    if not (energy == energy) or energy == float("inf"):
        print("the FDTD run went unstable", file=sys.stderr)
        return None
    value = nn * args.steps / secs / 1e6
    per = 2 if two else 1  # steps per launch of the sweep kernel
    launch_ms = ms[0] / max(timed, 1) * per
    # algorithmic bytes of the dominant (sweep) kernel per launch: E/H/coef/psi traffic only
    core_bpn = W.bytes_per_node_step(sim, None, core=True)
    alg = core_bpn * nn * per
    achieved = alg / (launch_ms * 1e-3) / 1e9
    # measured HBM bytes per launch of the sweep from the committed rocprofv3 PMC summary of this
    # workload (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel)
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"k_march_{args.workload}_traffic.json")
    if args.engine == "fused" and os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f)
        if t.get("steps_per_launch") == per:
            traffic = t.get("traffic_bytes_per_launch")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "alg_bytes_per_launch": round(alg), "steps_per_launch": per,
            "kernel": ("k_march2_mix or k_march2: two steps per launch over the interior, z-CPML and x/y-CPML "
                       "blocks (one mixed launch from 16 blocks per CU or below 8 M nodes, else one per class)"
                       if two else "k_march sweep (interior, z-CPML, CPML launches)" if args.engine == "fused"
                       else "k_update_h + k_update_e"),
            "kernel_ms": round(launch_ms, 4), "alg_bytes_per_node_step": round(core_bpn, 3),
            "fix_ms_per_step": round(ms[1] / max(timed, 1), 4), "observe_ms_per_step": round(ms[2] / max(timed, 1), 4)}
    if traffic:
        roof["traffic_gbs"] = round(traffic / (launch_ms * 1e-3) / 1e9, 1)
        roof["traffic_frac"] = round(roof["traffic_gbs"] / HBM_PEAK_GBS, 4)
        roof["traffic_over_alg"] = round(traffic / alg, 3)
    return {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(secs * 1e3 / args.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded materials and ports)",
        "config": dict(config, engine=args.engine, nodes=nn, bytes_per_node_step=round(bpn, 3),
                       hbm_frac_whole_step=round(bpn * nn * args.steps / secs / 1e9 / HBM_PEAK_GBS, 4),
                       build_s=round(build_s, 2), gpu=gpu.name),
        "roofline": roof,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="layered", choices=["layered", "boxes", "plain", "sweep", "adjoint"])
    ap.add_argument("--legs", default="sweep,adjoint",
                    help="layered: the legs added to the line (sweep, adjoint; 'none' for the headline alone)")
    ap.add_argument("--fd-cells", type=int, default=None,
                    help="adjoint: design cells checked by central differences (32 alone, 4 as a leg)")
    ap.add_argument("--launch", type=int, default=0,
                    help="sweep: two-step passes launched per pass (1) or persistently (2); 0 = autotuned")
    ap.add_argument("--graph-chunk", type=int, default=0,
                    help="sweep: steps per captured hipGraph (0 = autotuned, 1 = direct launches)")
    ap.add_argument("--engine", default="fused", choices=["fused", "split"])
    ap.add_argument("--timed-steps", type=int, default=50, help="steps timed per kernel with HIP events")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); multi-GPU halo exchange is not "
              "part of this build", file=sys.stderr)
        return 2
    if args.workload == "sweep":
        return sweep(args)
    if args.workload == "adjoint":
        args.fd_cells = 32 if args.fd_cells is None else args.fd_cells
        return adjoint(args)
    line = single(args)
    if line is None:
        return 1
    ok = True
    legs = [] if args.workload != "layered" or args.legs == "none" else args.legs.split(",")
    if "sweep" in legs:
        obj, _, same = sweep_leg(args)
        line["sweep"] = obj
        ok &= same
    if "adjoint" in legs:
        obj, _ = adjoint_leg(args, 4 if args.fd_cells is None else args.fd_cells)
        line["adjoint"] = obj
    if not args.no_cpu_baseline:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(args.workload)
    print(json.dumps(line), flush=True)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
