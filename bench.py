"""Throughput benchmark: Kalman track fit on one MI355X.

Headline workload (batched-fit-uneven, the batched kernel's own configuration
and the largest single-GPU one): 2,000,000 seeded toy-MC tracks in one batch,
25 barrel layers r = 3..110 cm plus 2 x 9 endcap disks |z| = 30..270 cm,
|eta| < 2.5, 8-30 hits per track, 5 % of layers missing, uniform Bz = 3.8 T,
x/X0 = 0.02 per layer with multiple scattering and energy loss, chi2
accumulated, FP64 covariance fallback on; FP32 state.  A "step" is one
complete fit of the batch (every track through all of its hits), inputs
resident in HBM, output written out of place.

Two forms of the same fit are timed on the same inputs:
  plain    -- mkf_fit_plain, one kernel launch per hit slot (state round-trips HBM)
  batched  -- mkf_fit_batched: one fused kernel over all slots, replayed from a
              cached hipGraph
The batch is packed in fit order (hit count, then endcap/barrel slot pattern;
mkfit_amd.fit_order), as a caller's packer would build it; both forms fit the
same packed batch.  --unordered packs it in generation order instead.
`value` is the batched throughput; the plain one and their ratio ride along.
The other workloads ride along in the same JSON line, each with its own
roofline and CPU baseline (skip them with --no-secondary):
  fit_at_scale   plain-fit-at-scale, 1,000,000 tracks, 25 barrel layers, one hit
                 per layer (--workload fit makes it the headline instead);
  find_ce / find_std / find_bh
                 clone-engine, standard combinatorial and best-hit finding of a
                 batch of 50 events (17 barrel layers, U(5k, 15k) seeds each,
                 noise 2 hits / cm^2), LayerOfHits build + one finder launch;
  backward_fit   the backward fit of every seed's best clone-engine candidate.

Prints one JSON line.  `--gpus N` with N > 1 exits: this build runs on one GPU.
"""
import argparse
import copy
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tracks fitted/sec on 1 MI355X, batched vs plain kernel, at ≤1e-5 rel param err"
# Algorithmic FP32 operations of one barrel track-layer step (propagate with
# material + chi2 + update) and of one endcap step (propagate to z with
# material + endcap chi2 + update), counted on the oracle restatement by
# oracle/flopcount.cc (+,-,*,/ = 1; sqrt, sin, cos, log = 1).
FLOP_PER_TRACK_LAYER = 1394
FLOP_PER_ENDCAP_STEP = 413 + 168
# Algorithmic HBM bytes per track-layer of the batched kernel: the hit (3 + 6
# floats) and its flag byte; per track: seed in (27 floats + charge), result out
# (27 floats + charge + chi2) and the hit count.
HIT_BYTES = 9 * 4 + 1
TRACK_BYTES = (27 + 1) * 4 + (27 + 1 + 1) * 4 + 4
# Plain kernel, per track-layer: state in and out (27 floats + charge + chi2 each
# way), the hit, its flag and the hit count.
PLAIN_BYTES = (27 + 1 + 1) * 4 * 2 + 9 * 4 + 1 + 4
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-tracks", type=int, default=None,
                    help="tracks of the fit workloads (default 2,000,000 uneven, 1,000,000 at scale)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle checker")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-secondary", action="store_true",
                    help="headline only: skip the fit-at-scale, finding and backward-fit sub-results")
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--shape", type=str, default=None,
                    help="block,blocks_per_cu,variant[,sort_tile] for the batched kernel (skips the autotuner)")
    ap.add_argument("--workload", default="fit-uneven",
                    choices=["fit-uneven", "fit", "find-ce", "find-bh", "find-std", "bkfit", "adjoint"],
                    help="fit-uneven (default, the headline): 2M tracks, barrel + 2x9 disks, 8-30 hits, 5%% "
                         "missing, FP64 fallback on; fit: plain-fit-at-scale, 1M tracks, 25 barrel layers; "
                         "find-ce / find-std / find-bh: clone-engine / standard / best-hit finding of --events "
                         "events; bkfit: backward fit of the clone engine's best candidates of --events events "
                         "(one untimed finding pass first); adjoint: gradient of the fit's chi2")
    ap.add_argument("--events", type=int, default=50,
                    help="events per step of the find workloads (all in one event of hits and one launch)")
    ap.add_argument("--unordered", action="store_true",
                    help="pack the fit batch in generation order instead of fit order (hit count, endcap pattern)")
    ap.add_argument("--probe", action="store_true",
                    help="counter-calibration run for rocprofv3 --pmc: one 256 MiB device copy, then "
                         "--steps batched and plain fits; no timing line")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Committed rocprofv3 summaries holding the copy-calibrated FETCH_SIZE + WRITE_SIZE of one batched launch.
TRAFFIC_PROFILES = {
    True: (os.path.join("profiles", "k_fit_batched-fit-uneven.md"), "k_fit batched uneven"),
    False: (os.path.join("profiles", "k_fit_plain-fit-at-scale.md"), "k_fit batched"),
}


def profiled_traffic(uneven):
    """HBM bytes per batched-fit launch, as measured by two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE, each
    scaled by a copy-kernel calibration) and recorded in the committed profile summary; None when absent."""
    import re
    path, label = TRAFFIC_PROFILES[uneven]
    try:
        text = open(os.path.join(ROOT, path)).read()
    except OSError:
        return None, path
    m = re.search(r"^traffic_bytes_per_launch\[%s\]: (\d+)" % re.escape(label), text, re.M)
    return (int(m.group(1)) if m else None), path


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch

    import mkfit_amd as mk

    dev = torch.device("cuda:0")
    torch.cuda.init()
    if args.workload in ("find-ce", "find-bh", "find-std"):
        res = bench_find(args, mk, torch, dev, args.workload[5:])
        res.pop("_out")
        print(json.dumps(res), flush=True)
        return
    if args.workload == "bkfit":
        a1 = copy.copy(args)
        a1.steps, a1.warmup, a1.no_cpu = 1, 0, True
        ce = bench_find(a1, mk, torch, dev, "ce")["_out"]
        print(json.dumps(bench_bkfit(args, mk, torch, dev, ce)), flush=True)
        return
    if args.workload == "adjoint":
        return bench_adjoint(args, mk, torch, dev)
    uneven = args.workload == "fit-uneven"
    res = bench_fit(args, mk, torch, dev, uneven)
    if res is None:  # --probe
        return
    if uneven and not args.no_secondary:
        a2 = copy.copy(args)
        a2.n_tracks, a2.shape = None, None
        sec = bench_fit(a2, mk, torch, dev, False)
        res["fit_at_scale"] = {k: sec[k] for k in ("value", "unit", "ms_per_step", "config", "plain",
                                                    "batched_over_plain", "parity", "roofline", "cpu_baseline")}
        keys = ("metric", "value", "unit", "ms_per_step", "config", "roofline", "cpu_baseline")
        finds = {}
        for kind in ("ce", "std", "bh"):
            r = bench_find(args, mk, torch, dev, kind)
            finds[kind] = r.pop("_out")
            res["find_" + kind] = {k: r[k] for k in keys}
        res["backward_fit"] = bench_bkfit(args, mk, torch, dev, finds["ce"])
    print(json.dumps(res), flush=True)


def bench_fit(args, mk, torch, dev, uneven):
    geo = mk.geometry(3 if uneven else 2)
    rl, xi = mk.material_grid(geo)
    if uneven:
        params = mk.default_toy_params(seed=args.seed, bfield=3.8, eta_max=2.5, hits_min=8, hits_max=30,
                                       missing_frac=0.05)
        n = args.n_tracks or 2_000_000
        n_slots = 30
    else:
        params = mk.default_toy_params(seed=args.seed, bfield=3.8)
        n = args.n_tracks or 1_000_000
        n_slots = geo.n_barrel
    t0 = time.time()
    ev = mk.generate(geo, params, n, n_slots)
    log("generated %d tracks in %.2f s; mean hits %.2f" % (n, time.time() - t0, ev["n_hits"].mean()))
    order_s = None
    if not args.unordered:
        # The batch is packed in fit order (hit count, then endcap/barrel slot pattern): both forms run on
        # this same packed batch; the host packing time is reported beside the kernels.
        t0 = time.perf_counter()
        ev, _ = mk.ordered_for_fit(ev)
        order_s = time.perf_counter() - t0
    pflags = mk.PF_APPLY_MATERIAL
    fflags = mk.FIT_CHI2 | (mk.FIT_FP64_FALLBACK if uneven else 0)
    ctx = mk.Context(bfield=params.bfield, rl_grid=rl, xi_grid=xi, device=0)
    batch = mk.FitBatch.from_numpy(ev, dev, pflags=pflags, fit_flags=fflags, out_of_place=True)
    plain_out = mk.FitBatch.from_numpy(ev, dev, pflags=pflags, fit_flags=fflags, out_of_place=True)
    n_track_layers = int(ev["n_hits"].sum())
    stream = torch.cuda.current_stream()

    tuned_ms = None
    if args.shape:
        ctx.set_batched_shape(*(int(x) for x in args.shape.split(",")))
    elif not args.no_autotune:
        t0 = time.time()
        shape, tuned_ms = ctx.autotune_batched(batch, reps=5)
        log("autotuned batched shape %s: %.3f ms (%.1f s)" % (shape, tuned_ms, time.time() - t0))
    shape = ctx.get_batched_shape()

    if args.probe:
        # A plain copy of known size: its FETCH_SIZE / WRITE_SIZE calibrate the counters.
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        for _ in range(args.steps):
            ctx.fit_batched(batch)
        ctx.fit_plain(plain_out)
        torch.cuda.synchronize()
        log("probe: copy %d bytes; %d batched fits and 1 plain fit of %d tracks, %d track-layers, shape %s" %
            (src.numel() * 4, args.steps, n, n_track_layers, shape))
        ctx.close()
        return None

    def timed(fn, steps, warmup):
        """Wall time of `steps` calls, and per call the GPU time between HIP events on the stream: the first
        pass brackets each call with its own events; the second runs the calls back to back between one
        pair of events (as the kernels run in the timed loop), whose mean is the per-launch duration."""
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        t_start = time.perf_counter()
        for e0, e1 in evs:
            e0.record(stream)
            fn()
            e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_start
        per = [e0.elapsed_time(e1) for e0, e1 in evs]
        a0, a1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a0.record(stream)
        for _ in range(steps):
            fn()
        a1.record(stream)
        torch.cuda.synchronize()
        return wall, per, a0.elapsed_time(a1) / steps

    wall_p, per_p, _ = timed(lambda: ctx.fit_plain(plain_out), args.steps, args.warmup)
    wall_b, per_b, launch_b = timed(lambda: ctx.fit_batched(batch), args.steps, args.warmup)
    ms_b = wall_b * 1e3 / args.steps
    ms_p = wall_p * 1e3 / args.steps
    tps_b = n / (ms_b * 1e-3)
    tps_p = n / (ms_p * 1e-3)
    # one graph per step (with a sort tile: ordering + fused kernel): HIP events on its stream, graphs back to back
    graph_ms = float(launch_b)
    order_ms = 0.0
    if shape[3] > 0:
        # the ordering kernel on its own (same kernel the graph replays), so the fused kernel's share is known
        order = torch.empty(n, dtype=torch.int32, device=dev)
        _, _, order_ms = timed(lambda: ctx.fit_order(batch, shape[3], order), args.steps, args.warmup)
    kern_b = graph_ms - order_ms
    log("plain %.3f ms/step (%.3g tracks/s); batched %.3f ms/step (%.3g tracks/s): graph %.3f ms = ordering %.3f "
        "+ fused kernel %.3f; batched/plain %.2f" % (ms_p, tps_p, ms_b, tps_b, graph_ms, order_ms, kern_b,
                                                     tps_b / tps_p))

    # The two GPU forms must agree bit for bit on the measured output.
    out_b = [batch.par.cpu().numpy(), batch.err.cpu().numpy(), batch.chi2.cpu().numpy(), batch.chg.cpu().numpy()]
    out_p = [plain_out.par.cpu().numpy(), plain_out.err.cpu().numpy(), plain_out.chi2.cpu().numpy(),
             plain_out.chg.cpu().numpy()]
    identical = all(np.array_equal(a, b, equal_nan=True) for a, b in zip(out_b, out_p))
    parity = {"batched_equals_plain": identical}

    cpu = None
    if not args.no_cpu:
        # CPU baseline leg: the oracle restatement, timed on the host, and (as the checker) a
        # sample of the measured GPU output against it, track by track.
        import oracle
        from oracle.parity import describe, fit_reference, flagged_parity
        ncpu = min(16, os.cpu_count() or 1)
        orc = oracle.Oracle(bfield=params.bfield, rl_grid=rl, xi_grid=xi)
        rng = np.random.default_rng(0)
        idx = np.sort(rng.choice(n, min(n, 20000), replace=False))
        sub = {k: np.ascontiguousarray(v[..., idx]) for k, v in ev.items()}
        o32, o64, chk = fit_reference(params.bfield, rl, xi, sub, pflags, fflags, n_threads=ncpu)
        r = flagged_parity((out_b[0][:, idx], out_b[1][:, idx], out_b[2][idx]), (o32[0], o32[1], o32[3]),
                           (o64[0], o64[1], o64[3]), chk)
        parity.update({"sample_tracks": int(len(idx)), "per_track": r,
                       "max_rel_par_err_vs_cpu": r["all_max"]["par"],
                       "max_rel_par_err_vs_cpu_unflagged": r["unflagged_max"]["par"],
                       "frac_tracks_within_1e-5": r["frac_tracks_within_tol"]})
        log("parity: batched==plain %s; %s" % (identical, describe(r)))
        best = 1e30
        for _ in range(args.cpu_reps):
            t = time.perf_counter()
            orc.fit(ev, pflags=pflags, fit_flags=fflags, n_threads=ncpu)
            best = min(best, time.perf_counter() - t)
        cpu = {"value": n / best, "unit": "tracks/s", "cores": ncpu, "kind": "port",
               "sample": "the full %d-track batch, best of %d passes, oracle FP32 Matriplex restatement "
                         "(NN=%d, %s), %d threads" % (n, args.cpu_reps, oracle.simd_width(), oracle.isa_name(),
                                                       ncpu)}
        log("cpu baseline %.3g tracks/s (%.2f s per pass)" % (n / best, best))

    flops = FLOP_PER_TRACK_LAYER * n_track_layers
    n_endcap = 0
    if uneven:   # endcap steps: propagation to z + endcap update (oracle/flopcount.cc)
        valid = np.arange(n_slots)[:, None] < ev["n_hits"][None, :]
        n_endcap = int(np.sum(valid & ((ev["hit_flags"] & 1) != 0)))
        flops = FLOP_PER_TRACK_LAYER * (n_track_layers - n_endcap) + FLOP_PER_ENDCAP_STEP * n_endcap
    achieved_tflops = flops / (kern_b * 1e-3) / 1e12
    alg_bytes_b = n_track_layers * HIT_BYTES + n * TRACK_BYTES
    traffic, traffic_src = profiled_traffic(uneven)
    if n != (2_000_000 if uneven else 1_000_000):
        traffic = None  # the profile holds the default size only
    res = {
        "metric": METRIC,
        "value": tps_b,
        "unit": "tracks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_b,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded toy MC (helix + multiple scattering + energy loss), seed %d" % args.seed,
        "config": {
            "workload": ("batched-fit-uneven: %d tracks, 25 barrel layers + 2x9 endcap disks |z|=30-270 cm, "
                         "|eta|<2.5, 8-30 hits per track, 5%% missing, Bz=3.8 T, x/X0=0.02, material on, chi2 "
                         "accumulated, FP64 covariance fallback" % n) if uneven else
                        ("plain-fit-at-scale: %d tracks, 25 barrel layers r=3-110 cm, |z|<=120 cm, Bz=3.8 T, "
                         "x/X0=0.02 per layer, material on, chi2 accumulated" % n),
            "n_tracks": n, "layers": n_slots, "track_layers": n_track_layers, "endcap_steps": n_endcap,
            "packing": ("fit order: tracks by hit count, then endcap/barrel slot pattern (host packing %.3f s, "
                        "outside the timed region; both forms fit this packed batch)" % order_s)
                       if order_s is not None else "generation order",
            "batched_shape": {"block": shape[0], "blocks_per_cu": shape[1], "variant": shape[2],
                              "sort_tile": shape[3]},
        },
        "plain": {"value": tps_p, "ms_per_step": ms_p, "unit": "tracks/s",
                  "achieved_GBps_algorithmic": n_track_layers * PLAIN_BYTES / (ms_p * 1e-3) / 1e9},
        "batched_over_plain": tps_b / tps_p,
        "parity": parity,
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / PEAK_FP32_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_fit (fused)", "kernel_ms": kern_b, "graph_ms": graph_ms,
                     "ordering_ms": order_ms,
                     "flop_per_track_layer": FLOP_PER_TRACK_LAYER, "flop_per_endcap_step": FLOP_PER_ENDCAP_STEP,
                     "algorithmic_bytes": alg_bytes_b,
                     "algorithmic_GBps": alg_bytes_b / (kern_b * 1e-3) / 1e9},
        "cpu_baseline": cpu,
    }
    ctx.close()
    return res


def _timed(torch, stream, fn, steps, warmup):
    for _ in range(warmup):
        fn(0)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t_start = time.perf_counter()
    for i, (e0, e1) in enumerate(evs):
        e0.record(stream)
        fn(i)
        e1.record(stream)
    torch.cuda.synchronize()
    return time.perf_counter() - t_start, [e0.elapsed_time(e1) for e0, e1 in evs]


# Algorithmic FP32 operations of the finding steps (oracle/flopcount.cc, barrel, no material): a propagation
# to a layer or to a hit, the chi2 of a hit, and the Kalman update of a state with its hit.
FIND_FLOP = {"propagate": 1031, "chi2": 65, "update": 282}


def find_events(args, mk):
    """The finding workload, generated once per run: --events toy events of the 17-layer barrel, U(5k, 15k)
    seeds each, noise 2 hits / cm^2, in one shared event of hits (event e owns layers e*17 ..)."""
    if getattr(args, "_find_events", None) is None:
        geo = mk.geometry(1)
        descs = mk.barrel_layer_descs(geo)
        rng = np.random.default_rng(args.seed)
        events = []
        t0 = time.time()
        for e in range(args.events):
            p = mk.default_toy_params(seed=args.seed + e, simulate_material=0)
            events.append(mk.generate_find_event(geo, p, int(rng.integers(5000, 15001)), noise_per_cm2=2.0))
        cat = mk.concat_find_events(events, len(descs))
        log("generated %d finding events in %.1f s; %d seeds, %d hits" % (len(events), time.time() - t0,
                                                                          cat["n_seeds"],
                                                                          int(cat["hit_offset"][-1])))
        args._find_events = (events, cat, descs)
    return args._find_events


def bench_find(args, mk, torch, dev, kind):
    """Track finding over a batch of events (find_events): clone engine (kind "ce") or standard combinatorial
    finding ("std") with up to 6 candidates and chi2 cut 30, propagating to each hit; or best-hit ("bh", chi2
    cut 15).  A step is the whole batch: the LayerOfHits build of every layer plus one finder launch over all
    seeds, inputs in HBM.  The roofline is the finder kernel's: its launch timed alone with HIP events on the
    stream, against the algorithmic FP32 operations of the work it does (propagations to the layers, window
    hits evaluated, updates), counted per seed by the oracle on the batch's first event (in the CPU baseline
    leg) and scaled to the batch's seeds.  Returns the result dict (the last GPU output in "_out")."""
    events, cat, descs = find_events(args, mk)
    L = len(descs)
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs * len(events))
    hits = torch.from_numpy(cat["hits_flat"]).to(dev)
    comb = kind in ("ce", "std")
    batch = (mk.CEBatch(cat, L, 3, device=dev, chi2_cut=30.0, max_cands=6, layer_base=cat["layer_base"]) if comb
             else mk.FindBatch(cat, L, 3, device=dev, chi2_cut=15.0, layer_base=cat["layer_base"]))
    finder = {"ce": ctx.find_tracks_clone_engine, "std": ctx.find_tracks_standard,
              "bh": ctx.find_tracks_best_hit}[kind]
    stream = torch.cuda.current_stream()
    kern = []

    def step(i):
        eoh.load(cat["hit_offset"], hits)
        if not comb:
            batch.reset()
        k0, k1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        k0.record(stream)
        finder(eoh, batch)
        k1.record(stream)
        kern.append((k0, k1))

    shape = None
    if comb and not args.no_autotune:
        # the finders' autotuner: slots per seed in a wavefront and waves per SIMD, timed on this batch
        eoh.load(cat["hit_offset"], hits)
        t0 = time.time()
        shape, tune_ms = ctx.autotune_find(eoh, batch, standard=kind == "std")
        log("find %s: autotuned shape (slots %d, waves/SIMD %d): %.2f ms per launch (%.1f s)" %
            (kind, shape[0], shape[1], tune_ms, time.time() - t0))
    wall, per = _timed(torch, stream, step, args.steps, args.warmup)
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in kern[args.warmup:]]))
    seeds = cat["n_seeds"] * args.steps
    ms = wall * 1e3 / args.steps
    prop_to_hit = comb
    unit_flop = {"propagate": FIND_FLOP["propagate"],
                 "evaluate": FIND_FLOP["chi2"] + (FIND_FLOP["propagate"] if prop_to_hit else 0),
                 "update": FIND_FLOP["update"] + (FIND_FLOP["propagate"] if prop_to_hit else 0)}
    name = {"ce": "clone-engine", "std": "standard combinatorial", "bh": "best-hit"}[kind]
    sl, wv = shape if shape else (6, 2)
    kname = {"ce": "k_find_clone_engine<%d, false, %d>" % (sl, wv), "std": "k_find_clone_engine<%d, true, %d>" % (sl, wv),
             "bh": "k_find_best_hit"}[kind]
    cpu = None
    roof = {"bound": "valu", "achieved": None, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": None,
            "traffic": None, "kernel": kname, "kernel_ms": kern_ms, "flop_per_unit": unit_flop}
    if not args.no_cpu:
        import oracle
        ncpu = min(16, os.cpu_count() or 1)
        ev = events[0]
        t = time.perf_counter()
        if comb:
            o = oracle.find_clone_engine(ev, descs, 3, chi2_cut=30.0, max_cands=6, n_threads=ncpu,
                                         standard=kind == "std", work=True)
        else:
            o = oracle.find_best_hit(ev, descs, 3, chi2_cut=15.0, n_threads=ncpu, work=True)
        dt = time.perf_counter() - t
        cpu = {"value": ev["n_seeds"] / dt, "unit": "seeds/s", "cores": ncpu, "kind": "port",
               "sample": "one event (%d seeds), oracle %s restatement incl. its LayerOfHits build, %d threads" %
                         (ev["n_seeds"], name, ncpu)}
        w = o["work"].astype(np.float64).sum(axis=1) / ev["n_seeds"]
        flop_seed = (w[0] * unit_flop["propagate"] + w[1] * unit_flop["evaluate"] + w[2] * unit_flop["update"])
        achieved = flop_seed * cat["n_seeds"] / (kern_ms * 1e-3) / 1e12
        roof.update({"achieved": achieved, "frac": achieved / PEAK_FP32_TFLOPS, "flop_per_seed": flop_seed,
                     "work_per_seed": {"propagations": w[0], "hits_evaluated": w[1], "updates": w[2]},
                     "work_source": "oracle work counts on the batch's first event (%d seeds), per seed, times "
                                    "the batch's %d seeds" % (ev["n_seeds"], cat["n_seeds"])})
    res = {"metric": "seeds processed/sec, %s track finding on 1 MI355X" % name,
           "value": seeds / wall, "unit": "seeds/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic: toy MC finding events (signal + uniform noise hits), seeds %d.." % args.seed,
           "config": {"workload": "find (%s), %d events per step in one launch: 17 barrel layers, U(5k,15k) seeds "
                                  "per event, noise 2/cm^2, %s" % (name, len(events),
                                                                   "up to 6 candidates, chi2 cut 30" if comb
                                                                   else "best hit, chi2 cut 15"),
                      "events": len(events), "seeds": cat["n_seeds"], "hits": int(cat["hit_offset"][-1]),
                      "find_shape": None if shape is None else {"slots": shape[0], "waves_per_simd": shape[1]}},
           "ms_per_event": ms / len(events), "roofline": roof, "cpu_baseline": cpu}
    log("find %s: %.3g seeds/s, %.2f ms per step, finder kernel %.2f ms%s" % (
        kind, res["value"], ms, kern_ms, "" if roof["frac"] is None else ", %.3f of FP32 peak" % roof["frac"]))
    res["_out"] = batch.numpy()
    ctx.close()
    return res


def bench_bkfit(args, mk, torch, dev, ce_out):
    """Backward fit of found candidates (mkf_backward_fit): the best clone-engine candidate of every seed of
    the finding batch (find_events), covariance x100, its hits in reverse layer order in blocks of 16
    candidates, then to the PCA.  A step is one backward fit of all candidates, inputs in HBM.  Roofline: the
    kernel (one launch) against 1322 FP32 operations per visited layer (propagation + chi2 + update) plus 413
    for the PCA step (oracle/flopcount.cc)."""
    events, cat, descs = find_events(args, mk)
    L = len(descs)
    keep = ce_out["n_cands"] > 0
    cand = {"par": ce_out["par"][0][:, keep], "err": ce_out["err"][0][:, keep], "chg": ce_out["chg"][0][keep],
            "hot_idx": ce_out["hot_idx"][0][:, keep], "score": ce_out["score"][0][keep]}
    base = cat["layer_base"][keep]
    n = int(keep.sum())
    ctx = mk.Context(bfield=3.8)
    eoh = mk.EventOfHits(ctx, descs * len(events))
    eoh.load(cat["hit_offset"], torch.from_numpy(cat["hits_flat"]).to(dev))
    b = mk.BkFitBatch(cand, device=dev, group=16, pflags=1, pca_pflags=1, layer_base=base)
    stream = torch.cuda.current_stream()
    wall, per = _timed(torch, stream, lambda i: ctx.backward_fit(eoh, b), args.steps, args.warmup)
    kern_ms = float(np.mean(per))
    # layers visited per candidate: those where some candidate of its block of 16 has a hit
    real = np.asarray(cand["hot_idx"]) >= 0
    G = 16
    pad = (-n) % G
    blk = np.concatenate([real, np.zeros((L, pad), bool)], axis=1).reshape(L, -1, G).any(axis=2)
    visited = int(blk.sum()) * G - pad * int(blk[:, -1].sum())      # the last block holds G - pad candidates
    flop = visited * (FIND_FLOP["propagate"] + 291) + n * 413
    achieved = flop / (kern_ms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        import oracle
        ncpu = min(16, os.cpu_count() or 1)
        ev = events[0]
        sel = np.nonzero(keep[:ev["n_seeds"]])[0]
        c0 = {k: np.ascontiguousarray(v[..., :len(sel)]) for k, v in cand.items()}
        t = time.perf_counter()
        oracle.backward_fit(ev, descs, c0, group=16, pflags=1, pca_pflags=1, bfield=3.8, n_threads=ncpu)
        dt = time.perf_counter() - t
        cpu = {"value": len(sel) / dt, "unit": "candidates/s", "cores": ncpu, "kind": "port",
               "sample": "the first event's %d candidates, oracle backward-fit restatement, %d threads" %
                         (len(sel), ncpu)}
    res = {"metric": "candidates backward-fitted/sec on 1 MI355X", "value": n * args.steps / wall,
           "unit": "candidates/s", "ms_per_step": wall * 1e3 / args.steps, "dtype": "f32",
           "config": {"workload": "backward fit: the best clone-engine candidate of each of %d seeds (%d events), "
                                  "17 barrel layers, blocks of 16, to the PCA" % (n, len(events)),
                      "candidates": n, "visited_layers": visited},
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "kernel": "k_backward_fit",
                        "kernel_ms": kern_ms, "flop_per_visited_layer": FIND_FLOP["propagate"] + 291,
                        "flop_pca": 413},
           "cpu_baseline": cpu}
    log("backward fit: %.3g candidates/s, kernel %.3f ms, %.3f of FP32 peak" % (res["value"], kern_ms,
                                                                              achieved / PEAK_FP32_TFLOPS))
    ctx.close()
    return res


def bench_adjoint(args, mk, torch, dev):
    """Gradient of the fit's total chi2 w.r.t. per-layer material and alignment: 200,000 tracks on the
    25-layer barrel, checkpoints every 5 layers.  A step is one forward fit + reverse sweep."""
    from mkfit_amd.alignment import make_event
    n = 200_000
    ev, geo, th = make_event(n, seed=args.seed)
    ctx = mk.Context(bfield=3.8)
    data = mk.AlignmentData(ctx, ev, dev, checkpoint_every=5)
    data64 = mk.AlignmentData(ctx, ev, dev, checkpoint_every=5, fp64=True)
    theta = torch.from_numpy(th.astype(np.float32)).to(dev)
    stream = torch.cuda.current_stream()
    wall, per = _timed(torch, stream, lambda i: data.run(theta), args.steps, args.warmup)
    wall64, _ = _timed(torch, stream, lambda i: data64.run(theta), args.steps, args.warmup)
    res = {"metric": "tracks/sec through fit + adjoint (d sum chi2 / d per-layer x/X0 and alignment) on 1 MI355X",
           "value": n * args.steps / wall, "unit": "tracks/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic: toy MC with per-layer x/X0 in [0.01, 0.05] and N(0, 50 um) offsets",
           "config": {"workload": "adjoint: %d tracks, 25 barrel layers, checkpoint every 5 layers" % n},
           "fp64": {"value": n * args.steps / wall64, "unit": "tracks/s", "ms_per_step": wall64 * 1e3 / args.steps,
                    "note": "the same model and gradient evaluated in FP64"},
           "cpu_baseline": None}
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
