"""Benchmark of the particle-history path on one MI355X.

Workload: the Al2O3-like target (Egap 8.8 eV, 4 VB + Al L + O K oscillator
sets, 3.95 g/cm^3, 100 A layer) hit by Au (Z 79) at 11 MeV/u, tallies on 1000
time nodes (0-1000 fs, dt 1 fs) x 100 log radial bins (1-1000 A).  Tables are
built once before timing (resident in HBM).  One step = trk_run_mc over
--iters-per-step Monte Carlo iterations (default 10^4, the configuration's
own count: whole ion histories with all their cascades), with distinct
iteration numbers so every step draws fresh random streams.  The metric is collision events per second (every SHI ionisation,
electron/hole elastic or inelastic collision, Auger or radiative decay and
photoabsorption counts once).

The value is the mixed-precision path (precision 1: FP32 sampling, FP64
state, ledger and tallies).  The same iterations, with the same random
streams, are then run on the FP64 path (precision 0, which replays the oracle
bit for bit): its events/s, the ratio and the error of every radial tally of
the mixed path against it are reported beside the value.

Also reported:
  roofline      the transport kernel's algorithmic bytes (per-event state
                read/write, spawned particles, table gathers, per-time-node
                snapshots of live particles) per launch over its mean launch
                time, HIP events on the library's stream, against 8 TB/s
  cpu_baseline  the oracle (C restatement of the reference's path) on the
                same tables and workload, iterations spread over the host
                cores as the reference's OpenMP loop does (and one core)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "trekis-3_amd"))

HBM_PEAK_GBS = 8000.0

# bytes of the algorithmic model (per event, per spawn, per lookup, per snapshot)
STATE_E, STATE_H = 72, 96          # 9 f64 / 11 f64 + 2 i32
SPAWN_PAIR, SPAWN_PHOT = STATE_E + STATE_H, STATE_E
GATHER = 16                        # two bracketing table entries
SNAP_E, SNAP_H = 48, 72


def transport_bytes(ev, nsh, snaps_e, snaps_h, snaps_p, phonon_rows):
    """Algorithmic bytes of the transport kernel's events (the ion kernel's
    SHI events excluded): (total, the part without the per-node snapshot
    terms)."""
    e_inel, e_el, h_inel, h_el, aug, rad, phot = (int(ev[i]) for i in range(1, 8))
    b = (e_inel + e_el) * 2 * STATE_E + (h_inel + h_el + aug + rad) * 2 * STATE_H + phot * 2 * STATE_E
    b += (e_inel + h_inel + aug + phot) * SPAWN_PAIR + rad * SPAWN_PHOT
    g = e_inel * (2 + nsh + 2) + e_el * (2 + (2 if phonon_rows else 0)) + h_inel * (2 + nsh + 2) \
        + h_el * (2 + (2 if phonon_rows else 0)) + (aug + rad) * nsh + phot * (1 + nsh)
    b += g * GATHER
    return b + snaps_e * SNAP_E + snaps_h * SNAP_H + snaps_p * SNAP_E, b


# radial tallies and the particle count that governs each: the error is the
# largest relative difference over the bins (radius x time node) at the final
# time node and three intermediate ones that hold at least 1% of the node's
# peak and where the FP64 path counts at least MIN_COUNTS particles of the
# tally's kind.  Out_Elat is taken as the lattice
# energy deposited up to the node, the quantity the reference writes out.
RADIAL = {"Out_ne": "Out_ne", "Out_Ee": "Out_ne", "Out_nh": "Out_nh", "Out_Eh": "Out_nh", "Out_Ehkin": "Out_nh",
          "Out_Elat": "Out_ne"}
PHOTON = {"Out_nphot": "Out_nphot", "Out_Ephot": "Out_nphot"}
MIN_COUNTS = 1000


def _tv(t, k):
    a = t[k]
    if k == "Out_Elat":
        return np.cumsum(a, axis=0)
    return a.sum(axis=(2, 3)) if a.ndim == 4 else a


# every other tally of do_Monte_Carlo: relative L1 difference over all its
# entries and the largest relative difference over the entries that hold at
# least 1% of the tally's peak (spectra and angle histograms are normalised
# per iteration, totals are per time node)
OTHER = ["Out_Ee_vs_E", "Out_Eh_vs_E", "Out_theta", "Out_theta_h", "Out_E_h", "Out_diff_coeff", "Out_tot_Ne",
         "Out_tot_Nphot", "Out_tot_E", "Out_E_e", "Out_E_phot", "Out_E_at"]


def other_errors(ref, mix):
    out = {}
    for k in OTHER:
        a, b = np.asarray(ref[k], dtype=np.float64), np.asarray(mix[k], dtype=np.float64)
        e = {"l1_rel": float(np.abs(b - a).sum() / np.abs(a).sum()) if np.abs(a).any() else 0.0}
        sel = np.abs(a) >= 1e-2 * np.abs(a).max() if np.abs(a).any() else np.zeros(a.shape, bool)
        sel &= a != 0.0
        if sel.any():
            e["max_rel_over_1pct_of_peak"] = float((np.abs(b - a)[sel] / np.abs(a)[sel]).max())
            e["entries"] = int(sel.sum())
        out[k] = e
    return out


def tally_errors(ref, mix, V):
    """Per-tally errors of `mix` against `ref` (see RADIAL): at the four
    nodes (max_rel), over every node for bins that also hold >= 1% of the
    tally's peak (all_nodes_max_rel), and the relative L1 difference over all
    bins.  Photon tallies are reported beside them, and every other tally
    through other_errors."""
    Nt = ref["Out_ne"].shape[0]
    nodes = sorted({Nt // 4, Nt // 2, (3 * Nt) // 4, Nt - 1})
    out, worst = {}, 0.0
    for k, kc in list(RADIAL.items()) + list(PHOTON.items()):
        a, b = _tv(ref, k), _tv(mix, k)
        counts = _tv(ref, kc) / V[None, :]
        e = {"l1_rel": float(np.abs(b - a).sum() / np.abs(a).sum()) if np.abs(a).any() else 0.0}
        an = a[nodes]
        peak = np.abs(an).max(axis=1, keepdims=True)
        sel = (counts[nodes] >= MIN_COUNTS) & (np.abs(an) >= 1e-2 * peak) & (an != 0.0)
        e["bins"] = int(sel.sum())
        if sel.any():
            e["max_rel"] = float((np.abs(b[nodes] - a[nodes])[sel] / np.abs(a[nodes])[sel]).max())
            if k in RADIAL:
                worst = max(worst, e["max_rel"])
        sel_all = (counts >= MIN_COUNTS) & (np.abs(a) >= 1e-2 * np.abs(a).max()) & (a != 0.0)
        if sel_all.any():
            e["all_nodes_max_rel"] = float((np.abs(b - a)[sel_all] / np.abs(a)[sel_all]).max())
            e["all_nodes_bins"] = int(sel_all.sum())
        out[k] = e
    te_a, te_b = float(ref["Out_tot_E"].sum()), float(mix["Out_tot_E"].sum())
    return worst, out, (abs(te_b - te_a) / abs(te_a)) if te_a else 0.0, nodes


def run_leg(tg, abi, args, precision, first_iter, quadrature=None):
    """Tables, warmup and the timed steps on one precision path; tallies of the
    timed steps accumulate in one set.  quadrature: the table quadrature
    (default: --quadrature for the mixed path, the reference's for FP64)."""
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(args.config)
    mat.prepare(lib)
    opt.gen_capacity = args.gen_capacity
    opt.precision = precision
    # the FP32 table quadrature serves the mixed path only; FP64 keeps the
    # reference's tables
    opt.quadrature = (args.quadrature if precision else 0) if quadrature is None else quadrature
    ctx = tg.Context(mat, ion, opt)
    t0 = time.perf_counter()
    ctx.build_tables()
    if args.config == "config5":
        # the large single track's synthetic photon attenuation tables
        tg.configs.apply_photon_tables(ctx, tg.configs.synthetic_photoabsorption(mat))
    t_tables = time.perf_counter() - t0
    name = "mixed" if precision else "fp64"
    print(f"[bench] {name}: tables built in {t_tables:.2f} s", file=sys.stderr, flush=True)
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    B = args.iters_per_step
    seed = run["seed"]
    warm_iter = first_iter + args.steps * B          # warmup draws streams the timed steps do not use
    for w in range(args.warmup):
        ctx.run_mc(B, seed, R, V, first_iter=warm_iter + w * B)
        print(f"[bench] {name}: warmup step {w} done", file=sys.stderr, flush=True)
    ctx.set_profiling(True)
    tal = ctx.new_tallies(len(R))
    ev = np.zeros(8, dtype=np.int64)
    t_total, host_ms = 0.0, []
    for k in range(args.steps):
        before = tal.n_events.copy()
        ts = time.perf_counter()
        ctx.run_mc(B, seed, R, V, tallies=tal, first_iter=first_iter + k * B)
        t_total += time.perf_counter() - ts
        step_ev = tal.n_events - before
        ev += step_ev
        print(f"[bench] {name}: step {k}: {int(step_ev.sum())} events", file=sys.stderr, flush=True)
        tr_s, tot_s, _ = ctx.timing()
        host_ms.append(1e3 * (tot_s - tr_s))
    ms_tr, n_tr = ctx.kernel_time(abi.K_TRANSPORT)
    ms_shi, n_shi = ctx.kernel_time(abi.K_SHI)
    ctx.set_profiling(False)
    # live particles at the time nodes (snapshot terms of the byte model)
    snaps = [float(tal["Out_tot_Ne"].sum()), float((tal["Out_nh"] / V[None, :, None, None]).sum()),
             float(tal["Out_tot_Nphot"].sum())]
    return {"ctx": ctx, "mat": mat, "ion": ion, "opt": opt, "run": run, "R": R, "V": V, "tal": tal,
            "ev": ev, "t_total": t_total, "host_ms": host_ms, "ms_tr": ms_tr, "n_tr": n_tr,
            "ms_shi": ms_shi, "n_shi": n_shi, "snaps": snaps, "t_tables": t_tables}


def run_batch_leg(tg, args, precision, n_impacts):
    """The batching workload: 4 material contexts, one ion context per impact,
    one Batch; the autotuner picks the launch shape on the first step's
    iterations, then --steps batches are timed (impact i of step k is
    iteration k*n + i)."""
    from trekis_gpu.batch import Batch, IonContext
    lib = tg.load_library()
    mats, opts, ions, run = tg.configs.config4_batch(n_impacts, precision)
    t0 = time.perf_counter()
    mctx = []
    for m, o in zip(mats, opts):
        m.prepare(lib)
        c = tg.Context(m, tg.abi.Ion.from_buffer_copy(ions[0][0]), o)
        c.build_tables()
        mctx.append(c)
    imps = []
    for ion, mi in ions:
        c = IonContext(mctx[mi], ion)
        c.build_tables()
        imps.append(c)
    t_tables = time.perf_counter() - t0
    name = "mixed" if precision else "fp64"
    print(f"[bench] batch {name}: {len(imps)} impacts on {len(mats)} materials, tables in {t_tables:.1f} s",
          file=sys.stderr, flush=True)
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mats[0].layer)
    b = Batch(imps, R, V, capacity=args.batch_capacity)
    n = len(imps)
    tune = None
    if args.tune:
        tune = b.tune(run["seed"], 10 ** 6)
        print(f"[bench] batch {name}: autotuner chose {tune}", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        b.run(run["seed"], (args.steps + w) * n, with_tallies=False)
    ev = np.zeros(8, dtype=np.int64)
    t_total, dev_ms, rounds = 0.0, 0.0, []
    tals = []
    fetch_s = 0.0
    for k in range(args.steps):
        # a step leaves every impact's finished tallies in HBM; the fetch of
        # all of them over PCIe (256 per-impact sets of the reference's
        # array shapes) is timed apart from the step
        ts = time.perf_counter()
        b.run(run["seed"], k * n, with_tallies=False)
        t_total += time.perf_counter() - ts
        ms, e, r = b.timing()
        dev_ms += ms
        ev += e
        rounds.append(r)
        if args.batch_tallies:
            ts = time.perf_counter()
            tals.append(b.fetch())
            fetch_s += time.perf_counter() - ts
        print(f"[bench] batch {name}: step {k}: {int(e.sum())} events, {ms:.1f} ms on the device, {r} rounds",
              file=sys.stderr, flush=True)
    return {"batch": b, "ev": ev, "t_total": t_total, "fetch_s": fetch_s, "dev_ms": dev_ms, "rounds": rounds, "tune": tune,
            "R": R, "seed": run["seed"],
            "tune_log": b.tune_log() if tune else None, "config": b.config(), "t_tables": t_tables, "tals": tals,
            "mats": mats, "ions": ions, "V": V, "impacts": imps, "mctx": mctx}


def batch_cpu_baseline(tg, args, leg):
    """The oracle on a bounded sample of the batch: the first impacts, one
    per host thread (impact i is iteration i of its own context, as on the
    GPU), each on that impact's GPU-built tables, in FP64, side by side."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = max(1, int(os.environ.get("OMP_NUM_THREADS", avail)))
    imps = leg["impacts"]
    sample = list(range(min(len(imps), threads)))
    R, V = leg["R"], leg["V"]
    orcs = []
    for i in sample:
        c = imps[i]
        o = Oracle(c.desc, c.ion, c.opt)
        o.inject_from(c)
        orcs.append(o)

    def one(j):
        t = tg.TallySet(orcs[j].Nt, len(R), imps[sample[j]].desc)
        t0 = time.perf_counter()
        orcs[j].run_mc(1, leg["seed"], R, V, t, first_iter=sample[j])
        return int(t.n_events.sum()), time.perf_counter() - t0

    ts = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(one, range(len(sample))))
    dt = time.perf_counter() - ts
    for o in orcs:
        o.close()
    evs = sum(e for e, _ in res)
    busy = sum(b for _, b in res)
    # histories differ ~100x in length, so a few impacts per thread leave
    # threads idle; the value is the all-core rate of a balanced schedule
    # (every thread busy, as the reference's dynamic OpenMP loop approaches
    # on a full batch): threads x events / summed per-history seconds
    return {"value": threads * evs / busy, "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"impacts 0-{len(sample) - 1} of the batch (one history each, one per host thread), FP64, "
                      f"on the GPU-built tables: {evs} events, {busy:.1f} s of summed per-history time; "
                      f"wall {dt:.1f} s ({evs / dt:.4g} events/s as run, threads idle after their history)",
            "single_core_value": evs / busy}


def main_batch(args):
    import trekis_gpu as tg
    n = args.impacts
    mixed = args.precision == "mixed"
    leg = run_batch_leg(tg, args, 1 if mixed else 0, n)
    total_ev = int(leg["ev"].sum())
    value = total_ev / leg["t_total"]
    compare = None
    if mixed and not args.no_compare:
        ref = run_batch_leg(tg, args, 0, n)
        ref_value = float(ref["ev"].sum()) / ref["t_total"]
        compare = {"fp64_value": ref_value, "speedup_vs_fp64": value / ref_value,
                   "fp64_autotuner_choice": ref["tune"]}
        if leg["tals"] and ref["tals"]:
            # radial tallies summed over each material's impacts and the steps
            worst, per = 0.0, {}
            for mi in range(len(leg["mats"])):
                idx = [i for i, (_, m) in enumerate(leg["ions"]) if m == mi]
                agg_m, agg_f = {}, {}
                for k in set(RADIAL) | set(RADIAL.values()) | set(PHOTON) | {"Out_tot_E"}:
                    agg_m[k] = sum(t[i][k] for t in leg["tals"] for i in idx)
                    agg_f[k] = sum(t[i][k] for t in ref["tals"] for i in idx)
                w, pm, de, nodes = tally_errors(agg_f, agg_m, leg["V"])
                worst = max(worst, w)
                per[f"material {mi} ({leg['mats'][mi].name})"] = {"max_radial_tally_rel_error": w, "tot_E_rel_diff": de,
                                             "tallies": {k: v for k, v in pm.items() if k in RADIAL}}
            compare["max_radial_tally_rel_error"] = worst
            compare["per_material"] = per
    cpu = batch_cpu_baseline(tg, args, leg) if args.cpu_iters > 0 else None
    line = {
        "metric": "collision events/s on 1 GPU, batched impacts, mixed vs FP64" if mixed
                  else "collision events/s on 1 GPU, batched impacts (FP64 path)",
        "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * leg["t_total"] / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 sampling / f64 state and tallies" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"{n} ion impacts (Z 18-92, 1-20 MeV/u) on 4 synthetic targets, one persistent "
                               f"launch per step, 0-1000 fs x 100 radial bins", "precision": args.precision,
                   "impacts_per_step": n, "events": total_ev,
                   "events_by_kind": dict(zip(["shi", "e_inel", "e_el", "h_inel", "h_el", "auger", "radiative",
                                               "photon"], leg["ev"].tolist())),
                   "device_ms_per_step": leg["dev_ms"] / args.steps, "rounds_per_launch": leg["rounds"],
                   "tally_fetch_ms_per_step": 1e3 * leg["fetch_s"] / args.steps if args.batch_tallies else None,
                   "pcie_inclusive_events_per_s": (total_ev / (leg["t_total"] + leg["fetch_s"])
                                                   if args.batch_tallies else None),
                   "launch": leg["config"], "autotuner": {"choice": leg["tune"], "timings": leg["tune_log"]},
                   "table_build_s": leg["t_tables"], "vs_fp64": compare},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters-per-step", type=int, default=None,
                    help="Monte Carlo iterations per step (default: config2 10^4, the configuration's own count; "
                         "config5 256)")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="path whose events/s is the value (mixed: also runs the FP64 path for the comparison)")
    ap.add_argument("--no-compare", action="store_true",
                    help="mixed only: skip the FP64 leg (no speed-up or tally error reported)")
    ap.add_argument("--cpu-iters", type=int, default=2,
                    help="oracle iterations per host thread for the CPU baseline (0 = skip)")
    ap.add_argument("--config", default="config2",
                    help="config2: Al2O3-like / Au 11 MeV/u; config4: the batch of ion impacts on 4 materials; "
                         "config5: the large single track (SiO2-like / U 20 MeV/u, 1 um, photons, phonons)")
    ap.add_argument("--impacts", type=int, default=256, help="config4: impacts per batch")
    ap.add_argument("--batch-capacity", type=int, default=1 << 22, help="config4: list capacity per kind and group")
    ap.add_argument("--tune", type=int, default=1, help="config4: run the autotuner first (1) or not (0)")
    ap.add_argument("--batch-tallies", type=int, default=1, help="config4: fetch every impact's tallies each step")
    ap.add_argument("--gen-capacity", type=int, default=1 << 27,
                    help="particles of one kind per round list (electrons, holes; photons get a quarter)")
    ap.add_argument("--quadrature", type=int, choices=[0, 1], default=0,
                    help="tables of the mixed leg: 0 the reference's FP64 quadrature, 1 the FP32 compensated "
                         "quadrature (2000 geometric transfer intervals x 1000 Simpson intervals in ln q)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the reference kernel "
                         "that scales HBM byte counters) and exit")
    args = ap.parse_args()
    if args.iters_per_step is None:
        args.iters_per_step = 256 if args.config == "config5" else 10000
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU split of "
              f"iterations is not built in this tree", file=sys.stderr)
        sys.exit(1)

    if args.config == "config4":
        main_batch(args)
        return
    import trekis_gpu as tg
    from trekis_gpu import abi
    mixed = args.precision == "mixed"
    main_leg = run_leg(tg, abi, args, 1 if mixed else 0, 0)
    mat, ion, opt, run = main_leg["mat"], main_leg["ion"], main_leg["opt"], main_leg["run"]
    R, V = main_leg["R"], main_leg["V"]
    B = args.iters_per_step
    ev, t_total = main_leg["ev"], main_leg["t_total"]
    total_ev = int(ev.sum())
    value = total_ev / t_total
    ms_tr, n_tr = main_leg["ms_tr"], main_leg["n_tr"]
    nsh = sum(mat.n_shells)
    b_tr, b_ev = transport_bytes(ev, nsh, *[int(round(s)) for s in main_leg["snaps"]], opt.kind_of_EMFP == 1)
    per_launch = b_tr / max(n_tr, 1)
    per_launch_ev = b_ev / max(n_tr, 1)
    mean_s = (ms_tr / max(n_tr, 1)) * 1e-3
    achieved = per_launch / mean_s / 1e9 if mean_s > 0 else 0.0
    achieved_moved = per_launch_ev / mean_s / 1e9 if mean_s > 0 else 0.0
    if achieved_moved > HBM_PEAK_GBS:
        raise SystemExit(f"roofline fraction above 1 ({achieved_moved:.0f} GB/s): the byte model is wrong")

    compare = None
    if mixed and not args.no_compare:
        ref = run_leg(tg, abi, args, 0, 0)
        worst, per, de, nodes = tally_errors(ref["tal"], main_leg["tal"], V)
        per.update(other_errors(ref["tal"], main_leg["tal"]))
        ref_value = float(ref["ev"].sum()) / ref["t_total"]
        compare = {"fp64_value": ref_value, "speedup_vs_fp64": value / ref_value,
                   "fp64_transport_mean_launch_ms": ref["ms_tr"] / max(ref["n_tr"], 1),
                   "max_radial_tally_rel_error": worst, "tot_E_rel_diff": de, "tallies": per,
                   "error_rule": "largest relative difference of Out_ne, Out_Ee, Out_nh, Out_Eh, Out_Ehkin and "
                                 f"cumulative Out_Elat at time nodes {nodes} (0-based) over the radial bins that "
                                 f"hold >= 1% of the node's peak and where the FP64 path counts >= {MIN_COUNTS} "
                                 "particles of the tally's kind; same iterations and random streams on both paths"}
        ref["ctx"].close()
        if args.quadrature:
            # the transport's own error: the FP64 path on the same FP32-quadrature
            # tables (the error above also holds the table difference, the
            # reference quadrature's left-node bias of 0.2-1.3 %)
            same = run_leg(tg, abi, args, 0, 0, quadrature=args.quadrature)
            w2, per2, de2, _ = tally_errors(same["tal"], main_leg["tal"], V)
            compare["max_radial_tally_rel_error_vs_reference_tables"] = compare["max_radial_tally_rel_error"]
            compare["tot_E_rel_diff_vs_reference_tables"] = compare["tot_E_rel_diff"]
            compare["max_radial_tally_rel_error"] = w2
            compare["tot_E_rel_diff"] = de2
            compare["tallies_vs_reference_tables"] = compare["tallies"]
            per2.update(other_errors(same["tal"], main_leg["tal"]))
            compare["tallies"] = per2
            compare["error_rule"] += ("; max_radial_tally_rel_error and tallies: against the FP64 path on the same "
                                      "FP32-quadrature tables; *_vs_reference_tables: against the FP64 path on the "
                                      "reference's quadrature (the speed-up's denominator)")
            same["ctx"].close()

    # measured HBM bytes per launch of the transport kernel: rocprofv3 --pmc
    # runs of this workload (FETCH_SIZE, then WRITE_SIZE, scaled by a plain
    # copy) committed under profiles/; counters cannot be read inside this run
    traffic, traffic_src = None, None
    tf = os.path.join(ROOT, "profiles", f"k_transport__al2o3_au11_{B}iter_{args.precision}_traffic.txt")
    if args.config == "config2" and os.path.exists(tf):
        kv = dict(l.strip().split("=", 1) for l in open(tf) if "=" in l and not l.startswith("#"))
        traffic = float(kv["traffic_bytes_per_launch"])
        traffic_src = os.path.relpath(tf, ROOT)

    cpu = None
    if args.cpu_iters > 0:
        # the oracle (C restatement of the reference's path) on the host cores:
        # iterations spread over threads like the reference's OpenMP loop, and
        # one iteration on one core
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        seed = run["seed"]
        orc = Oracle(mat, ion, opt)
        orc.inject_from(main_leg["ctx"])
        # every host core this process may use (the box's share when
        # OMP_NUM_THREADS says so), as the reference's Num_th = -1 does
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, int(os.environ.get("OMP_NUM_THREADS", avail)))
        model = ""
        try:
            model = next((l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")), "")
        except OSError:
            pass
        n_all = args.cpu_iters * threads
        ts = time.perf_counter()
        c = orc.run_mc_threads(n_all, seed, R, V, tg.TallySet(orc.Nt, len(R), mat), threads)
        dt = time.perf_counter() - ts
        ts1 = time.perf_counter()
        c1 = orc.run_mc(1, seed, R, V, tg.TallySet(orc.Nt, len(R), mat))
        dt1 = time.perf_counter() - ts1
        cpu = {"value": float(c.n_events.sum()) / dt, "unit": "events/s", "cores": threads, "kind": "port",
               "sample": f"{n_all} iterations of the same workload on the GPU-built tables, FP64, OpenMP over "
                         f"iterations on {threads} threads: {int(c.n_events.sum())} events in {dt:.1f} s; "
                         f"one iteration on one core: {int(c1.n_events.sum())} events in {dt1:.1f} s",
               "single_core_value": float(c1.n_events.sum()) / dt1,
               "cpu_model": model, "os_cpu_count": os.cpu_count(), "cores_available": avail}

    kname = "trk::mx::k_transport" if mixed else "trk::fp64::k_transport"
    line = {
        "metric": "collision events/s on 1 GPU, mixed vs FP64, at <=1% radial-tally error" if mixed
                  else "collision events/s on 1 GPU (FP64 path)",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sampling / f64 state and tallies" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": (f"{mat.name} target, Au 11 MeV/u ion, 0-1000 fs x 100 radial bins, "
                                if args.config != "config5" else
                                f"{mat.name} target, U 20 MeV/u ion over a 1 um track, phonon quasi-elastic scattering, "
                                f"photons on synthetic attenuation tables, 0-1000 fs x 200 radial bins, ")
                               + f"{B} iterations per step"
                               + (", mixed leg on FP32-quadrature tables" if mixed and args.quadrature else ""),
                   "precision": args.precision, "iters_per_step": B,
                   "table_quadrature": "fp32 compensated" if mixed and args.quadrature else "reference fp64",
                   "events": total_ev, "events_by_kind": dict(zip(abi.EVENT_NAMES, ev.tolist())),
                   "table_build_s": main_leg["t_tables"], "ion_init_kernel_ms": main_leg["ms_shi"],
                   "ion_init_launches": main_leg["n_shi"],
                   "outside_batches_ms_per_step": sum(main_leg["host_ms"]) / len(main_leg["host_ms"]),
                   "vs_fp64": compare},
        # achieved: the algorithmic bytes the kernel moves per launch (state
        # read and written per event, spawned particles, table gathers) over
        # its mean launch time.  A model of the reference's loop also counts a
        # per-time-node snapshot of every live particle, which the difference-array tallies
        # never make; that figure is reported beside it, not as the fraction.
        "roofline": {"bound": "hbm", "achieved": achieved_moved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved_moved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "measured_limiter": "issue and latency of dependent table lookups under lane divergence "
                                         "(HBM is not what bounds the kernel: see traffic_frac)",
                     "kernel": kname, "launches": n_tr, "mean_launch_ms": ms_tr / max(n_tr, 1),
                     "alg_bytes_per_launch": per_launch_ev,
                     "model_bytes_per_launch_with_snapshots": per_launch,
                     "model_GBs_with_snapshots": achieved,
                     "traffic_GBs": (traffic / mean_s / 1e9) if (traffic and mean_s > 0) else None,
                     "traffic_frac": (traffic / mean_s / 1e9 / HBM_PEAK_GBS) if (traffic and mean_s > 0) else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
