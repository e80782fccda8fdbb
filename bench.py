"""Benchmark of the particle-history path on one MI355X.

Workload: the Al2O3-like target (Egap 8.8 eV, 4 VB + Al L + O K oscillator
sets, 3.95 g/cm^3, 100 A layer) hit by Au (Z 79) at 11 MeV/u, FP64, tallies on
1000 time nodes (0-1000 fs, dt 1 fs) x 100 log radial bins (1-1000 A).  Tables
are built once before timing (resident in HBM).  One step = trk_run_mc over
--iters-per-step Monte Carlo iterations (whole ion histories with all their
cascades), with distinct iteration numbers so every step draws fresh random
streams.  The metric is collision events per second (every SHI ionisation,
electron/hole elastic or inelastic collision, Auger or radiative decay and
photoabsorption counts once).

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
    SHI events excluded)."""
    e_inel, e_el, h_inel, h_el, aug, rad, phot = (int(ev[i]) for i in range(1, 8))
    b = (e_inel + e_el) * 2 * STATE_E + (h_inel + h_el + aug + rad) * 2 * STATE_H + phot * 2 * STATE_E
    b += (e_inel + h_inel + aug + phot) * SPAWN_PAIR + rad * SPAWN_PHOT
    g = e_inel * (2 + nsh + 2) + e_el * (2 + (2 if phonon_rows else 0)) + h_inel * (2 + nsh + 2) \
        + h_el * (2 + (2 if phonon_rows else 0)) + (aug + rad) * nsh + phot * (1 + nsh)
    b += g * GATHER
    b += snaps_e * SNAP_E + snaps_h * SNAP_H + snaps_p * SNAP_E
    return b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iters-per-step", type=int, default=4096)
    ap.add_argument("--cpu-iters", type=int, default=2,
                    help="oracle iterations per host thread for the CPU baseline (0 = skip)")
    ap.add_argument("--config", default="config2")
    ap.add_argument("--gen-capacity", type=int, default=1 << 27,
                    help="particles of one kind per round list (electrons, holes; photons get a quarter)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the reference kernel "
                         "that scales HBM byte counters) and exit")
    args = ap.parse_args()
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

    import trekis_gpu as tg
    from trekis_gpu import abi
    lib = tg.load_library()
    mat, ion, opt, run = tg.config_material(args.config)
    mat.prepare(lib)
    opt.gen_capacity = args.gen_capacity
    ctx = tg.Context(mat, ion, opt)
    t0 = time.perf_counter()
    ctx.build_tables()
    t_tables = time.perf_counter() - t0
    print(f"[bench] tables built in {t_tables:.2f} s", file=sys.stderr, flush=True)
    R, V = tg.radial_grid(run["NR"], run["Rmin"], run["Rmax"], mat.layer)
    B = args.iters_per_step
    seed = run["seed"]
    it = 0
    for w in range(args.warmup):
        ctx.run_mc(B, seed, R, V, first_iter=it)
        it += B
        print(f"[bench] warmup step {w} done", file=sys.stderr, flush=True)

    ctx.set_profiling(True)
    ev = np.zeros(8, dtype=np.int64)
    snaps = [0.0, 0.0, 0.0]
    t_total = 0.0
    host_ms = []
    for k in range(args.steps):
        t = ctx.new_tallies(len(R))
        ts = time.perf_counter()
        ctx.run_mc(B, seed, R, V, tallies=t, first_iter=it)
        t_total += time.perf_counter() - ts
        it += B
        ev += t.n_events
        snaps[0] += float(t["Out_tot_Ne"].sum())
        snaps[1] += float((t["Out_nh"] / V[None, :, None, None]).sum())
        snaps[2] += float(t["Out_tot_Nphot"].sum())
        print(f"[bench] step {k}: {int(t.n_events.sum())} events", file=sys.stderr, flush=True)
        tr_s, tot_s, _ = ctx.timing()
        host_ms.append(1e3 * (tot_s - tr_s))
    ms_tr, n_tr = ctx.kernel_time(abi.K_TRANSPORT)
    ms_shi, n_shi = ctx.kernel_time(abi.K_SHI)
    ctx.set_profiling(False)

    total_ev = int(ev.sum())
    value = total_ev / t_total
    nsh = sum(mat.n_shells)
    b_tr = transport_bytes(ev, nsh, *[int(round(s)) for s in snaps], opt.kind_of_EMFP == 1)
    per_launch = b_tr / max(n_tr, 1)
    mean_s = (ms_tr / max(n_tr, 1)) * 1e-3
    achieved = per_launch / mean_s / 1e9 if mean_s > 0 else 0.0

    # measured HBM bytes per launch of the transport kernel (FETCH_SIZE and
    # WRITE_SIZE passes, scaled by a plain copy), for this workload if taken
    traffic = None
    tf = os.path.join(ROOT, "profiles", f"k_transport__al2o3_au11_{B}iter_traffic.txt")
    if args.config == "config2" and os.path.exists(tf):
        kv = dict(l.strip().split("=", 1) for l in open(tf) if "=" in l and not l.startswith("#"))
        traffic = float(kv["traffic_bytes_per_launch"])

    cpu = None
    if args.cpu_iters > 0:
        # the oracle (C restatement of the reference's path) on the host cores:
        # iterations spread over threads like the reference's OpenMP loop, and
        # one iteration on one core
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        orc = Oracle(mat, ion, opt)
        orc.inject_from(ctx)
        threads = max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 16))
        n_all = args.cpu_iters * threads
        ts = time.perf_counter()
        c = orc.run_mc_threads(n_all, seed, R, V, tg.TallySet(orc.Nt, len(R), mat), threads)
        dt = time.perf_counter() - ts
        ts1 = time.perf_counter()
        c1 = orc.run_mc(1, seed, R, V, tg.TallySet(orc.Nt, len(R), mat))
        dt1 = time.perf_counter() - ts1
        cpu = {"value": float(c.n_events.sum()) / dt, "unit": "events/s", "cores": threads, "kind": "port",
               "sample": f"{n_all} iterations of the same workload on the GPU-built tables, OpenMP over "
                         f"iterations on {threads} threads: {int(c.n_events.sum())} events in {dt:.1f} s; "
                         f"one iteration on one core: {int(c1.n_events.sum())} events in {dt1:.1f} s",
               "single_core_value": float(c1.n_events.sum()) / dt1}

    line = {
        "metric": "collision events/s on 1 GPU (FP64 path)",
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{mat.name} target, Au 11 MeV/u ion, 0-1000 fs x 100 radial bins, "
                               f"{B} iterations per step", "iters_per_step": B, "events": total_ev,
                   "events_by_kind": dict(zip(abi.EVENT_NAMES, ev.tolist())), "table_build_s": t_tables,
                   "ion_init_kernel_ms": ms_shi, "ion_init_launches": n_shi,
                   "outside_batches_ms_per_step": sum(host_ms) / len(host_ms)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_transport", "launches": n_tr, "mean_launch_ms": ms_tr / max(n_tr, 1),
                     "alg_bytes_per_launch": per_launch,
                     "traffic_GBs": (traffic / mean_s / 1e9) if (traffic and mean_s > 0) else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
