"""Benchmark of the Barnes-Hut connectivity update on one MI355X.

Workload (default): 16M neurons uniform in a 25,000 um cube, 80/20 E/I, vacant axons ~ Poisson(2),
vacant excitatory/inhibitory dendrites ~ Poisson(1.5), theta 0.3, sigma 750 um, leaf capacity 1.
A step is one plasticity update on the device-resident state: vacancies restored to the workload's
values (device-to-device copy), leaf fill + upward pass, target search, conflict resolution and
commit.  The FP32-mixed and the FP64 paths run the same seeded input; `value` is the FP32-mixed
whole-step throughput in target searches per second.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "target searches/s @16M neurons, FP32-mixed vs FP64; ms per plasticity step"
HBM_PEAK_GBS = 8000.0
B_NODE = {"fp64": 40, "fp32_mixed": 24}   # bytes read per acceptance test (record + count/link)
B_IO = 32                                  # source position + output ids per search
# Simulation workloads: Normal(5, 2) input alone drives the Izhikevich neurons to about 0.0117 spikes
# per step, so calcium settles near 0.117 and never reaches the reference's axon growth threshold
# 0.4, and with no axons no synapse forms to add input; the "growing" thresholds and target let
# elements grow and connect, the "reference" ones are the configuration as stated.
SIM_THRESHOLDS = {"growing": ((0.02, 0.01, 0.0), 0.2), "reference": ((0.4, 0.1, 0.0), 0.7)}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_flags(kernel):
    from relearn_amd import RL_FLAG_EXACT_ORDER, RL_FLAG_WALK_ALWAYS, RL_FLAG_WAVE_PER_SOURCE

    return {"default": 0, "wave": RL_FLAG_WAVE_PER_SOURCE, "walk": RL_FLAG_WALK_ALWAYS,
            "lane": RL_FLAG_EXACT_ORDER}[kernel]


def walk_counts(net, precision, seed, d_vac, flags, tuning):
    """Node steps and table slots of one FP32-mixed group-walk search, from the walk's counting
    instantiation (RL_FLAG_COUNT_WALK) on a context of its own, outside every timed region: the counts
    depend on the tree, the positions and the vacancies only, which every step restores.  (0, 0) when
    this precision and kernel choice do not run the walk."""
    from relearn_amd import RL_FLAG_COUNT_WALK, RL_PREC_FP32_MIXED, BarnesHutHIP

    if precision != RL_PREC_FP32_MIXED or flags & 3:
        return 0.0, 0.0, 0.0
    g = BarnesHutHIP(net.box_min, net.box_max, theta=0.3, sigma=750.0, seed=seed, precision=precision,
                     flags=flags | RL_FLAG_COUNT_WALK)
    g.init(net.xyz, net.signal)
    if tuning:
        g.set_tuning(tuning["search_waves_per_cu"], tuning["search_chunk"], tuning.get("walk_min_sources", 0))
    g.prepare_update_connectivity_device(*d_vac)
    g.find_target_neurons(0)
    st = g.last_search_stats()
    g.close()
    log(f"group walk: {st['walk_steps']} node steps, {st['walk_slots']} table slots, "
        f"{st['walk_table_bytes']} table bytes stored per search")
    return float(st["walk_steps"]), float(st["walk_slots"]), float(st["walk_table_bytes"])


def run_precision(net, precision, steps, warmup, seed, d_vac, flags=0):
    from relearn_amd import BarnesHutHIP

    g = BarnesHutHIP(net.box_min, net.box_max, theta=0.3, sigma=750.0, seed=seed, precision=precision, flags=flags)
    t0 = time.perf_counter()
    g.init(net.xyz, net.signal)
    build_s = time.perf_counter() - t0
    m, depth = g.tree_info()
    log(f"precision {precision}: tree {m} nodes, depth {depth}, build {build_s:.3f} s")
    tuning = None
    if not (flags & 1):
        # search launch tuning for this workload (results do not depend on it)
        g.prepare_update_connectivity_device(*d_vac)
        tuning = g.autotune(999)
        log(f"precision {precision}: autotuned {tuning}")

    def one_step(step):
        t_a = time.perf_counter()
        g.prepare_update_connectivity_device(*d_vac)
        t_b = time.perf_counter()
        g.find_target_neurons(step)
        t_c = time.perf_counter()
        st = g.last_search_stats()
        created = g.commit(step)
        t_d = time.perf_counter()
        return st, created, (t_b - t_a, t_c - t_b, t_d - t_c)

    for w in range(warmup):
        one_step(1000 + w)
    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    recs = []
    for k in range(steps):
        recs.append(one_step(k))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    searches = sum(r[0]["searches"] for r in recs)
    tests = sum(r[0]["node_tests"] for r in recs)
    tests_ref = sum(r[0]["node_tests_ref"] for r in recs)
    kernel_ms = [r[0]["ms_search"] for r in recs]
    dominant_ms = [r[0]["ms_search_kernel"] for r in recs]
    # the same update captured as one HIP graph and replayed (vacancies restored before each replay,
    # outside the timed replay)
    slots = int(net.vac_axon.astype(np.uint64).sum())
    g.prepare_update_connectivity_device(*d_vac)
    g.graph_create(0, slots)
    for w in range(warmup):
        g.prepare_update_connectivity_device(*d_vac)
        g.graph_launch(0, 1000 + w)
    graph_ms = []
    for k in range(steps):
        g.prepare_update_connectivity_device(*d_vac)
        graph_ms.append(g.graph_launch(0, k)["ms_total"])
    g.graph_destroy()
    # pick results of step 0 for the precision comparison
    g.prepare_update_connectivity_device(*d_vac)
    g.find_target_neurons(0)
    targets, _, counts = g.get_search_slots(slots)
    created0 = g.commit(0)
    # in-degree of every neuron after that update, and the level-1 candidate lists of a fixed
    # sample of sources (probabilities of this precision) for the precision comparison
    _, ve1, vi1 = g.get_vacancies()
    indeg = (net.vac_exc.astype(np.int64) - ve1) + (net.vac_inh.astype(np.int64) - vi1)
    g.prepare_update_connectivity_device(*d_vac)
    sample = np.random.default_rng(seed + 7).choice(net.n, min(1000, net.n), replace=False)
    cand = g.candidate_lists(sample, capacity=16384)
    g.close()
    walk_steps, walk_slots, walk_bytes = walk_counts(net, precision, seed, d_vac, flags, tuning)
    return {
        "build_s": build_s,
        "nodes": m,
        "depth": depth,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "searches_per_step": searches / steps,
        "searches_per_s": searches / wall,
        "search_ms": float(np.mean(kernel_ms)),
        "dominant_kernel_ms": float(np.mean(dominant_ms)),
        "search_phase_searches_per_s": searches / (1e-3 * sum(kernel_ms)),
        "node_tests_per_step": tests / steps,
        "walk_steps_per_step": walk_steps,
        "walk_slots_per_step": walk_slots,
        "walk_table_bytes_per_step": walk_bytes,
        "node_tests_ref_per_step": tests_ref / steps,
        "phase_s": [float(np.mean([r[2][i] for r in recs])) for i in range(3)],
        "graph_ms_per_update": float(np.mean(graph_ms)),
        "tuning": tuning,
        "requests_per_step": float(np.mean([r[0]["requests"] for r in recs])),
        "created_per_step": float(np.mean([r[1] for r in recs])),
        "targets": targets,
        "counts": counts,
        "indeg": indeg,
        "cand": cand,
        "created0": created0,
    }


def run_ensemble(args):
    """Batched plasticity updates of a parameter-sweep ensemble (one context, one launch per phase)
    next to the same networks updated one context at a time."""
    import torch

    from relearn_amd import RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP
    from relearn_amd._lib import network_seed
    from relearn_amd.workloads import parameter_sweep

    t0 = time.perf_counter()
    ens = parameter_sweep(args.networks, seed=args.seed)
    G, n = ens.num_networks, ens.n
    log(f"ensemble: {G} networks, {n} neurons, generated in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    d_arrays = [torch.from_numpy(a.view(np.int32)).to(dev) for a in (ens.vac_axon, ens.vac_exc, ens.vac_inh)]
    torch.cuda.synchronize()
    d_vac = [int(a.data_ptr()) for a in d_arrays]
    lo, hi = ens.boxes[:, :3].min(0), ens.boxes[:, 3:].max(0)
    out = {"metric": "ensemble target searches/s (parameter sweep, one context, one launch per phase)",
           "unit": "searches/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
           "data": "synthetic",
           "config": {"workload": f"{G} networks, sizes log-uniform [2e3, 2e5] ({n} neurons), cube side "
                                  "1000 (n/1000)^(1/3) um, theta 0.3, sigma 750 um, one plasticity update per step",
                      "networks": G, "neurons": n, "seed": args.seed}}
    for name, prec in (("fp32_mixed", RL_PREC_FP32_MIXED), ("fp64", RL_PREC_FP64)):
        g = BarnesHutHIP(lo, hi, theta=0.3, sigma=750.0, seed=args.seed + 1, precision=prec,
                         flags=kernel_flags(args.kernel))
        t0 = time.perf_counter()
        g.init_networks(ens.offsets, ens.boxes, ens.xyz, ens.signal)
        build_s = time.perf_counter() - t0
        g.prepare_update_connectivity_device(*d_vac)
        tuning = g.autotune(999)
        for w in range(args.warmup):
            g.prepare_update_connectivity_device(*d_vac)
            g.find_target_neurons(1000 + w)
            g.commit(1000 + w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        searches = 0
        for k in range(args.steps):
            g.prepare_update_connectivity_device(*d_vac)
            g.find_target_neurons(k)
            searches += g.last_search_stats()["searches"]
            g.commit(k)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # the same update of all networks captured as one graph
        g.prepare_update_connectivity_device(*d_vac)
        g.graph_create(0, int(ens.vac_axon.astype(np.uint64).sum()))
        graph_ms = []
        for k in range(args.warmup + args.steps):
            g.prepare_update_connectivity_device(*d_vac)
            ms = g.graph_launch(0, k)["ms_total"]
            if k >= args.warmup:
                graph_ms.append(ms)
        g.close()
        out[name] = {"build_s": build_s, "ms_per_update": 1e3 * wall / args.steps, "searches_per_s": searches / wall,
                     "graph_ms_per_update": float(np.mean(graph_ms)), "tuning": tuning}
        log(f"ensemble {name}: {1e3 * wall / args.steps:.1f} ms per update of all networks, "
            f"{searches / wall:.4g} searches/s, graph replay {np.mean(graph_ms):.1f} ms, forest build {build_s:.2f} s")
    # the same networks, one context each, updated one after another (unbatched)
    if not args.no_unbatched:
        ctxs = []
        t0 = time.perf_counter()
        for k in range(G):
            a, b = int(ens.offsets[k]), int(ens.offsets[k + 1])
            c = BarnesHutHIP(ens.boxes[k, :3], ens.boxes[k, 3:], theta=0.3, sigma=750.0,
                             seed=network_seed(args.seed + 1, k), precision=RL_PREC_FP32_MIXED)
            c.init(ens.xyz[a:b], ens.signal[a:b])
            ctxs.append((c, a))
        init_s = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        searches = 0
        for step in range(args.steps):
            for c, a in ctxs:
                c.prepare_update_connectivity_device(*[p + 4 * a for p in d_vac])
                c.find_target_neurons(step)
                searches += c.last_search_stats()["searches"]
                c.commit(step)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        for c, _ in ctxs:
            c.close()
        out["fp32_mixed_unbatched"] = {"init_s": init_s, "ms_per_update": 1e3 * wall / args.steps,
                                       "searches_per_s": searches / wall}
        out["batching_speedup"] = out["fp32_mixed_unbatched"]["ms_per_update"] / out["fp32_mixed"]["ms_per_update"]
        log(f"unbatched fp32_mixed: {1e3 * wall / args.steps:.1f} ms per update of all networks")
    out["value"] = out["fp32_mixed"]["searches_per_s"]
    out["ms_per_step"] = out["fp32_mixed"]["ms_per_update"]
    out["dtype"] = "f32-mixed"
    out["vs_baseline"] = None
    out["scaling"] = "weak"
    print(json.dumps(out), flush=True)
    return 0


def run_sim(args):
    """Simulation periods replayed as HIP graphs: `--period` neuron steps (Izhikevich, calcium, growth
    curves, counter-based Normal(5, 2) input), the element commit and the connectivity update, one
    graph replay per period.  Runs FP32-mixed and FP64 with the same seed and reports the mean
    calcium and synapses-per-neuron trajectories of both."""
    import torch

    from relearn_amd import RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP, BHError, StepParams
    from relearn_amd.workloads import parameter_sweep, uniform_cube

    t0 = time.perf_counter()
    if args.workload == "sim-ensemble":
        ens = parameter_sweep(args.networks, seed=args.seed)
        n, G = ens.n, ens.num_networks
        lo, hi = ens.boxes[:, :3].min(0), ens.boxes[:, 3:].max(0)
        desc = (f"{G} networks, sizes log-uniform [2e3, 2e5] ({n} neurons), growth rates log-uniform "
                f"[{5e-5 * args.nu_scale:g}, {5e-4 * args.nu_scale:g}] per network")
    else:
        net = uniform_cube(args.sim_neurons, args.sim_side, seed=args.seed)
        n, G = net.n, 1
        lo, hi = net.box_min, net.box_max
        desc = f"{n} neurons uniform in a {args.sim_side:g} um cube, growth rate {1e-4 * args.nu_scale:g}"
    log(f"sim workload: {desc}, generated in {time.perf_counter() - t0:.1f} s")
    periods, period = args.sim_periods, args.period
    sim_eta, sim_target = SIM_THRESHOLDS[args.thresholds]
    every = max(1, periods // 10)
    out = {"metric": "simulated neuron steps/s (plasticity every period steps, one graph replay per period)",
           "unit": "neuron-steps/s", "n_gpus": 1, "steps": periods, "warmup": 0, "higher_is_better": True,
           "data": "synthetic",
           "config": {"workload": f"{desc}; Izhikevich + calcium + Gaussian growth, input Normal(5, 2), "
                                  f"growth thresholds eta {sim_eta}, target calcium {sim_target} ({args.thresholds}), theta 0.3, "
                                  f"sigma 750 um, {periods} periods of {period} steps, elements start at 0, "
                                  f"equally weighted synaptic input (conductance {args.conductance:g}) over the "
                                  "network the updates build",
                      "neurons": n, "networks": G, "periods": periods, "period_steps": period, "seed": args.seed}}
    traj = {}
    for name, prec in (("fp32_mixed", RL_PREC_FP32_MIXED), ("fp64", RL_PREC_FP64)):
        g = BarnesHutHIP(lo, hi, theta=0.3, sigma=750.0, seed=args.seed + 1, precision=prec,
                         flags=kernel_flags(args.kernel))
        if args.workload == "sim-ensemble":
            g.init_networks(ens.offsets, ens.boxes, ens.xyz, ens.signal)
        else:
            g.init(net.xyz, net.signal)
        # search launch tuning on a representative update (Poisson vacancies), then the run starts
        # from zero elements as configured
        vr = np.random.default_rng(args.seed + 3)
        g.prepare_update_connectivity(vr.poisson(2.0, n).astype(np.uint32), vr.poisson(1.5, n).astype(np.uint32),
                                      vr.poisson(1.5, n).astype(np.uint32))
        tuning = g.autotune(999)
        zero = np.zeros(n, np.uint32)
        g.prepare_update_connectivity(zero, zero, zero)
        p = StepParams.izhikevich(input_mode=1, nu=1e-4 * args.nu_scale)
        for t in range(3):
            p.eta[t] = sim_eta[t]
        # synaptic input over the device network graph, which every update extends
        p.synaptic_input = 1
        p.synapse_conductance = args.conductance
        g.net_init(24 * n)
        g.neurons_init(p, np.full(n, sim_target), x=np.full(n, -65.0), u=np.full(n, -13.0))
        if args.workload == "sim-ensemble":
            g.neurons_set_network_growth(np.repeat(ens.nu * args.nu_scale, 3))
        cap = 4 * n
        g.graph_create(period, cap)
        rows, created, deleted, regrown, net_regrown = [], 0, 0, 0, 0
        torch.cuda.synchronize()
        wall = 0.0   # replays only (each waits for its period); checkpoint copies are not timed
        for per in range(periods):
            t_run = time.perf_counter()
            try:
                st = g.graph_launch(per * period, per)
                wall += time.perf_counter() - t_run
            except BHError as e:
                if e.code != -7 or "max_vacant_axons" not in str(e):
                    raise
                # more vacant axons than the captured capacity: the replay's update changed nothing,
                # so the eager update of the same index completes the period; then grow the graph
                wall += time.perf_counter() - t_run
                need = int(g.get_vacancies()[0].sum())
                t_run = time.perf_counter()
                st = g.update_connectivity(per)
                wall += time.perf_counter() - t_run
                cap = max(2 * cap, need + need // 4)
                log(f"sim {name}: period {per}: {need} vacant axons over max_vacant_axons, captured again for {cap}")
                g.graph_create(period, cap)
                regrown += 1
            if st.get("net_regrown"):
                net_regrown += 1
                log(f"sim {name}: period {per}: network graph grown to {st['net_capacity']} entries")
            created += st["created"]
            deleted += st.get("deleted", 0)
            if (per + 1) % every == 0 or per + 1 == periods:
                s = g.neurons_get()
                rows.append({"step": (per + 1) * period, "mean_ca": float(s["ca"].mean()),
                             "synapses_per_neuron": float(s["connected"][0].sum()) / n,
                             "vacant_axons": int(g.get_vacancies()[0].sum())})
        edges = int(g.net_get_edges()[0][-1])
        g.close()
        out[name] = {"ms_per_period": 1e3 * wall / periods, "neuron_steps_per_s": n * period * periods / wall,
                     "created": created, "deleted": deleted, "graph_regrown": regrown,
                     "net_regrown": net_regrown, "edges": edges,
                     "tuning": tuning, "trajectory": rows}
        traj[name] = rows
        log(f"sim {name}: {1e3 * wall / periods:.2f} ms per period of {period} steps, "
            f"{n * period * periods / wall:.4g} neuron-steps/s, {created} synapses created")
    a, b = traj["fp64"], traj["fp32_mixed"]
    out["fp32_vs_fp64"] = {
        "max_rel_diff_mean_ca": max(abs(x["mean_ca"] - y["mean_ca"]) / max(abs(x["mean_ca"]), 1e-300) for x, y in zip(a, b)),
        "max_rel_diff_synapses_per_neuron": max(abs(x["synapses_per_neuron"] - y["synapses_per_neuron"])
                                                / max(x["synapses_per_neuron"], 1e-300) for x, y in zip(a, b)),
        "speedup": out["fp64"]["ms_per_period"] / out["fp32_mixed"]["ms_per_period"],
    }
    out["value"] = out["fp32_mixed"]["neuron_steps_per_s"]
    out["ms_per_step"] = out["fp32_mixed"]["ms_per_period"] / period
    out["dtype"] = "f64 step, f32-mixed search"
    out["vs_baseline"] = None
    out["scaling"] = "weak"
    print(json.dumps(out), flush=True)
    return 0


def copy_calibration(mib):
    """Runs a plain device-to-device copy of `mib` MiB three times (reference kernel for the HBM counters)."""
    import torch

    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}), flush=True)
    return 0


def precision_comparison(a, b):
    """Probability L-infinity on level-1 candidate lists (probabilities normalised by the FP64 list
    total) and the total-variation distance of the in-degree histograms, FP64 run a vs FP32-mixed b."""
    linf, rel, same, longest = 0.0, 0.0, True, 0
    for (n64, p64, c64), (n32, p32, c32) in zip(a["cand"], b["cand"]):
        longest = max(longest, c64)
        if c64 != c32 or not np.array_equal(n64, n32):
            same = False
            continue
        if len(p64) == 0:
            continue
        tot = float(p64.sum())
        if tot > 0:
            linf = max(linf, float(np.abs(p32 - p64).max()) / tot)
        # candidates that can matter to a pick (FP32 exp2 flushes probabilities below ~1e-38 to 0)
        pos = p64 > 1e-12 * tot
        if pos.any():
            rel = max(rel, float((np.abs(p32 - p64)[pos] / p64[pos]).max()))
    h64, h32 = np.bincount(a["indeg"]), np.bincount(b["indeg"])
    m = max(len(h64), len(h32))
    h64 = np.pad(h64, (0, m - len(h64)))
    h32 = np.pad(h32, (0, m - len(h32)))
    return {"probability_linf": linf, "probability_max_rel_err_above_1e-12": rel, "candidate_lists_identical": same,
            "candidate_sample": len(a["cand"]), "longest_candidate_list": longest,
            "in_degree_tv_distance": 0.5 * float(np.abs(h32 - h64).sum()) / max(1, int(h64.sum()))}


def read_traffic(kernel, precision, neurons):
    """HBM bytes per search launch measured by rocprofv3 (FETCH_SIZE + WRITE_SIZE, copy-calibrated),
    from the committed summary profiles/<kernel>_<precision>_<neurons>_traffic.txt, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{precision}_{neurons}_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip()
    try:
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def read_counters(kernel, precision_label):
    """Limiter figures of the search kernel from the committed rocprofv3 counter summary
    profiles/<kernel>_<precision>_16M_counters.md (lines '- <name>: <value>' under 'Derived'), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{precision_label}_16M_counters.md")
    if not os.path.exists(path):
        return None
    out = {"profile": os.path.relpath(path, ROOT)}
    keys = {"limiter": "resource", "VALU busy": "valu_busy", "SALU busy": "salu_busy",
            "texture data busy": "td_busy", "L2 hit rate": "l2_hit_rate",
            "VALU instructions per active source":
            "valu_insts_per_source", "SALU instructions per active source": "salu_insts_per_source",
            "commit": "profiled_commit"}
    with open(path) as f:
        for line in f:
            if not line.startswith("- ") or ":" not in line:
                continue
            k, v = line[2:].split(":", 1)
            for prefix, name in keys.items():
                if k.strip().startswith(prefix):
                    v = v.strip()
                    try:
                        out[name] = float(v)
                    except ValueError:
                        out[name] = v
    return out


def host_threads():
    """Host cores this process may use: OMP_NUM_THREADS when set (the GPU host's per-GPU share), else the
    CPUs of the process's affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(net, sample, threads, seed):
    from oracle import pyoracle

    t0 = time.perf_counter()
    t = pyoracle.Tree(net.xyz, net.box_min, net.box_max)
    t.update(net.vac_exc, net.vac_inh, threads=threads)
    prep = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    cand = np.nonzero(net.vac_axon)[0]
    src = np.sort(rng.choice(cand, min(sample, len(cand)), replace=False))
    r = t.find_targets_faithful(net.signal, net.vac_axon, src, theta=0.3, sigma=750.0, seed=seed, threads=threads)
    t.close()
    return r, prep, len(src)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--neurons", type=int, default=16_000_000)
    ap.add_argument("--side", type=float, default=25_000.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=20_000, help="source neurons searched by the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32_mixed", "fp64_plain"], default=None)
    ap.add_argument("--no-plain", action="store_true", help="skip the plain FP64 (one lane per source) run")
    ap.add_argument("--kernel", choices=["default", "wave", "walk", "lane"], default="default",
                    help="search kernel: default (FP32-mixed: the group walk from 192 x 64 active sources per CU, "
                         "one wavefront per source below; FP64: one wavefront per source), wave (one wavefront per "
                         "source always), walk (FP32-mixed: the group walk always), lane (one lane per source in "
                         "list order)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (reference kernel for rocprofv3 byte counters)")
    ap.add_argument("--workload", choices=["single", "ensemble", "sim-ensemble", "sim-single"], default="single",
                    help="single: one 16M-neuron network (default); ensemble: batched parameter-sweep networks; "
                         "sim-ensemble / sim-single: simulation periods replayed as graphs")
    ap.add_argument("--sim-periods", type=int, default=100, help="simulation: plasticity periods")
    ap.add_argument("--period", type=int, default=100, help="simulation: neuron steps per plasticity update")
    ap.add_argument("--sim-neurons", type=int, default=4_000_000, help="sim-single: neurons")
    ap.add_argument("--sim-side", type=float, default=16_000.0, help="sim-single: cube side in micrometres")
    ap.add_argument("--nu-scale", type=float, default=10.0, help="simulation: growth rates x this factor")
    ap.add_argument("--conductance", type=float, default=0.03,
                    help="simulation: synapse conductance (the reference's default 0.03)")
    ap.add_argument("--thresholds", choices=["growing", "reference"], default="growing",
                    help="simulation: growth thresholds eta 0.02/0.01/0 with target calcium 0.2 (axons grow), "
                         "or the reference's 0.4/0.1/0 with target 0.7 (calcium stays below the axons' threshold)")
    ap.add_argument("--networks", type=int, default=256, help="networks of the ensemble workload")
    ap.add_argument("--no-unbatched", action="store_true", help="ensemble: skip the one-context-per-network run")
    args = ap.parse_args(argv)
    # OpenMP threads of the CPU baseline stay on neighbouring cores (read when the runtime loads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    if args.calibrate_copy:
        return copy_calibration(args.calibrate_copy)

    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the sharded design is documented, not built)", file=sys.stderr)
        return 2

    if args.workload == "ensemble":
        return run_ensemble(args)
    if args.workload.startswith("sim-"):
        return run_sim(args)

    from relearn_amd import RL_FLAG_EXACT_ORDER, RL_FLAG_WAVE_PER_SOURCE, RL_PREC_FP32_MIXED, RL_PREC_FP64
    from relearn_amd.workloads import uniform_cube
    import torch

    t0 = time.perf_counter()
    net = uniform_cube(args.neurons, args.side, seed=args.seed)
    log(f"workload {args.neurons} neurons in {args.side} um cube generated in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    d_arrays = [torch.from_numpy(a.view(np.int32)).to(dev) for a in (net.vac_axon, net.vac_exc, net.vac_inh)]
    torch.cuda.synchronize()
    d_vac = [int(a.data_ptr()) for a in d_arrays]

    res = {}
    kflags = kernel_flags(args.kernel)
    # fp64_plain: the FP64 path with one lane per source and the CPU path's list order (the plain
    # FP64 GPU method the FP32-mixed path is compared with)
    precs = [("fp32_mixed", RL_PREC_FP32_MIXED, kflags), ("fp64", RL_PREC_FP64, kflags)]
    if not args.no_plain:
        precs.append(("fp64_plain", RL_PREC_FP64, RL_FLAG_EXACT_ORDER))
    if args.only:
        precs = [p for p in precs if p[0] == args.only]
    for name, prec, flags in precs:
        res[name] = run_precision(net, prec, args.steps, args.warmup, args.seed + 1, d_vac, flags=flags)
        r = res[name]
        log(f"{name}: {r['searches_per_s']:.4g} searches/s whole step, {r['ms_per_step']:.1f} ms/step, "
            f"search {r['search_ms']:.1f} ms (kernel {r['dominant_kernel_ms']:.1f} ms), phases {['%.3f' % x for x in r['phase_s']]} s, "
            f"graph replay {r['graph_ms_per_update']:.1f} ms")

    main_name = "fp32_mixed" if "fp32_mixed" in res else ("fp64" if "fp64" in res else "fp64_plain")
    main_r = res[main_name]
    b_node = B_NODE["fp32_mixed" if main_name == "fp32_mixed" else "fp64"]
    # algorithmic bytes of one launch.  The wave-per-source and per-lane kernels read one node record for
    # every acceptance test they execute; the group walk reads one node record per step (shared by the
    # up to 64 sources of the wavefront) and stores, for each expanded node with an expanded child, the
    # running sums of the sources that expanded it and its child entries (counted by the walk's counting
    # instantiation).  Every search reads its source and writes its result.
    if main_r["walk_steps_per_step"] > 0:
        bytes_per_launch = (main_r["walk_steps_per_step"] * b_node + main_r["walk_table_bytes_per_step"]
                            + main_r["searches_per_step"] * B_IO)
        bytes_model = (f"walk steps x {b_node} B (node record, read once for the wavefront) + table bytes stored "
                       f"(8 B per running sum kept, 32 B per child-entry row) + searches x {B_IO} B per launch")
    else:
        bytes_per_launch = main_r["node_tests_per_step"] * b_node + main_r["searches_per_step"] * B_IO
        bytes_model = f"executed acceptance tests x {b_node} B + searches x {B_IO} B per launch"
    kernel_s = 1e-3 * main_r["dominant_kernel_ms"]
    achieved = bytes_per_launch / kernel_s / 1e9
    if args.kernel == "lane" or main_name == "fp64_plain":
        kname = "k_search"
    elif main_name == "fp32_mixed" and main_r["walk_steps_per_step"] > 0:
        kname = "k_walk"   # the group walk ran (its counting run counted steps)
    else:
        kname = "k_level1"
    traffic = read_traffic(kname, main_name, args.neurons)
    counters = read_counters(kname, main_name) if args.neurons == 16_000_000 else None
    out = {
        "metric": METRIC,
        "value": main_r["searches_per_s"],
        "unit": "searches/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed" if main_name == "fp32_mixed" else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.neurons} neurons uniform in a {args.side:g} um cube, 80/20 E/I, vacant axons "
                        "Poisson(2), vacant dendrites Poisson(1.5) per type, theta 0.3, sigma 750 um, one "
                        "plasticity update per step",
            "neurons": args.neurons,
            "side_um": args.side,
            "seed": args.seed,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": kname,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_model": bytes_model,
            "kernel_ms": main_r["dominant_kernel_ms"],
            "node_tests_per_launch": main_r["node_tests_per_step"],
            "walk_steps_per_launch": main_r["walk_steps_per_step"],
            "walk_slots_per_launch": main_r["walk_slots_per_step"],
            "walk_table_bytes_per_launch": main_r["walk_table_bytes_per_step"],
            # the same executed tests priced at one node record each (the wave-per-source kernels' demand)
            "achieved_per_test_records": (main_r["node_tests_per_step"] * b_node + main_r["searches_per_step"] * B_IO)
            / kernel_s / 1e9,
            "searches_per_launch": main_r["searches_per_step"],
            # the reference's traversal re-derives the level-1 list for every axon: its test count is
            # larger than the kernel's, which derives it once per neuron
            "achieved_reference_tests": (main_r["node_tests_ref_per_step"] * b_node + main_r["searches_per_step"] * B_IO)
            / kernel_s / 1e9,
            "hbm_physical_gbs": None if traffic is None else traffic / kernel_s / 1e9,
            "hbm_physical_frac": None if traffic is None else traffic / kernel_s / 1e9 / HBM_PEAK_GBS,
            # what limits the kernel, from its rocprofv3 counters: instruction issue, not HBM
            "limiter": counters,
        },
    }
    for name, r in res.items():
        out[name] = {k: v for k, v in r.items() if k not in ("targets", "counts", "indeg", "cand")}
    if "fp64" in res and "fp32_mixed" in res:
        a, b = res["fp64"], res["fp32_mixed"]
        valid = a["targets"] != 0xFFFFFFFF
        div = float(np.mean(a["targets"][valid] != b["targets"][valid]))
        out["fp32_vs_fp64"] = {
            "speedup_whole_step": a["ms_per_step"] / b["ms_per_step"],
            "speedup_search": a["search_ms"] / b["search_ms"],
            "pick_divergence": div,
            "requests_fp64": int(a["counts"].sum()),
            "requests_fp32": int(b["counts"].sum()),
            "mean_synapses_per_neuron_fp64": a["created0"] / args.neurons,
            "mean_synapses_per_neuron_fp32": b["created0"] / args.neurons,
            "synapses_rel_diff": abs(b["created0"] - a["created0"]) / max(1, a["created0"]),
        }
        out["fp32_vs_fp64"].update(precision_comparison(a, b))
    if "fp64_plain" in res and "fp32_mixed" in res:
        a, b = res["fp64_plain"], res["fp32_mixed"]
        valid = a["targets"] != 0xFFFFFFFF
        out["fp32_vs_fp64_plain"] = {
            "speedup_whole_step": a["ms_per_step"] / b["ms_per_step"],
            "speedup_search": a["search_ms"] / b["search_ms"],
            "pick_divergence": float(np.mean(a["targets"][valid] != b["targets"][valid])),
            "synapses_rel_diff": abs(b["created0"] - a["created0"]) / max(1, a["created0"]),
        }
    if not args.no_cpu:
        threads = args.cpu_threads or host_threads()
        r, prep, ns = cpu_baseline(net, args.cpu_sample, threads, args.seed + 1)
        sps = r["searches"] / r["seconds"]
        r1, _, ns1 = cpu_baseline(net, max(1, args.cpu_sample // 8), 1, args.seed + 2)
        sps1 = r1["searches"] / r1["seconds"]
        out["cpu_baseline"] = {
            "value": sps,
            "unit": "searches/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{r['searches']} searches of {ns} random source neurons on the full "
                      f"{args.neurons}-neuron pointer octree (mt19937 stream, omp parallel for + critical "
                      f"appends, OMP_PROC_BIND=close); tree build + upward pass {prep:.1f} s not in the rate",
            "cpu_model": cpu_model(),
            "host_cpus_visible": os.cpu_count(),
            "single_thread_value": sps1,
            "single_thread_sample": f"{r1['searches']} searches of {ns1} source neurons, 1 thread",
            "search_ms_per_step_extrapolated": 1e3 * main_r["searches_per_step"] / sps,
            "gpu_speedup": out["value"] / sps,
        }
        log(f"cpu baseline: {sps:.4g} searches/s on {threads} threads ({r['searches']} searches), "
            f"{sps1:.4g} on 1 thread")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
