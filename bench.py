#!/usr/bin/env python3
"""Swendsen-Wang throughput of the MI355X engine (site-updates/s).

Default workload: the batched uniform scan -- L=32, p=0.05, 40 temperatures
in linspace(1.0, 3.0) x 32 disorder realizations = 1280 lattices, one
workgroup per lattice with everything in LDS, 2000 thermalisation + 20000
measured sweeps (10 bins) with the moments accumulated on the device.  A step
is one sweep of every lattice of the batch; the default times the whole
22000-sweep run.  Disorder is synthetic (seeded numpy draws, as the reference
generates it), spins start hot.

Printed as one JSON line, with
  roofline      achieved algorithmic HBM bytes of the dominant kernel
                (4 B per site-update: spin read + write, J_h + J_v read) per
                average launch duration, vs MI355X's 8 TB/s; with the measured
                HBM traffic and, as valu_issue, the share of the VALU issue
                cycles the kernel's profiled instruction count fills (its
                actual bound);
  plain         the one-lattice-at-a-time tiled path at the same L;
  cpu_baseline  the oracle's C restatement of the reference sweep on all host
                cores (one lattice per thread), on a bounded sample (~10 s);
  L256          the metric's second point, the same recipe at L=256 (1280
                lattices of 65536 sites, 200 + 2000 sweeps): multi-tile
                union-find in both the batched and the plain mode.

    python bench.py [--gpus 1] [--steps 22000] [--warmup 200] [--second-L 256]
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rbim-fk-percolation_amd"))

PEAK_HBM_GBS = 8000.0
# VALU issue capacity: 256 CUs x 4 SIMDs at the 2.4 GHz peak engine clock; a
# wave64 VALU instruction holds its SIMD for 4 cycles
VALU_SIMD_CYCLES_PER_S = 1024 * 2.4e9
ALGO_BYTES_PER_SITE_UPDATE = 4


def scan_specs(L, p, temperatures, R, seed):
    from rbim.engine import LatticeSpec, engine_seed
    from rbim.disorder import generate_disorder
    from rbim.simulation import scan_lattices

    disorder_seeds, run_seeds = np.random.SeedSequence(seed).spawn(2)
    rng = np.random.default_rng(disorder_seeds)
    realizations = [generate_disorder(L, p, rng) for _ in range(R)]
    return scan_lattices(L, realizations, temperatures, engine_seed(run_seeds))


def run_schedule(K):
    """(n_therm, n_bins, sweeps_per_bin) with n_therm + n_bins * spb == K, ~1/11
    thermalisation and up to 10 bins (the default scan's shape at K = 22000)."""
    measured = max(1, K - K // 11)
    n_bins = min(10, measured)
    spb = measured // n_bins
    return K - n_bins * spb, n_bins, spb


def gpu_batched(specs, K, W, engine, sweeps_per_launch=0, threads=0, streams=0, tile=0):
    from rbim.engine import Batch

    n_therm, n_bins, spb = run_schedule(K)
    if W > 0:
        wn, wb, ws = run_schedule(W)
        with Batch(specs, n_therm=wn, n_bins=wb, sweeps_per_bin=ws, engine=engine,
                   sweeps_per_launch=sweeps_per_launch, threads=threads, streams=streams,
                   tile=tile) as warm:
            warm.run_timed(W)
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=engine,
               sweeps_per_launch=sweeps_per_launch, threads=threads, streams=streams,
               tile=tile) as batch:
        t0 = time.perf_counter()
        ms = batch.run_timed(K)
        wall = time.perf_counter() - t0
        info = batch.info()
        acc = batch.moments()
    sites = sum(s.L * s.L for s in specs)
    return {
        "ms": ms,
        "wall_s": wall,
        "site_updates": sites * K,
        "launches": int(info.kernel_launches),
        "threads": int(info.threads),
        "sweeps_per_launch": int(info.sweeps_per_launch),
        "checksum": int(acc[:, :, 17].sum()),
    }


def gpu_plain(spec, K):
    from rbim.engine import Batch

    n_therm, n_bins, spb = run_schedule(K)
    with Batch([spec], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=1) as warm:
        warm.run_timed(min(K, 64))
    with Batch([spec], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=1) as batch:
        ms = batch.run_timed(K)
    return spec.L * spec.L * K / (ms / 1e3)


def profiled_fields(L, n_lat, sweeps_per_launch, threads):
    """The `key: integer` lines of the committed rocprofv3 summary of this
    benchmark's batch (HBM traffic from FETCH_SIZE / WRITE_SIZE calibrated
    against a plain device copy; SQ VALU instructions per wavefront-sweep of
    the dominant kernel), or {} when the configuration is not the profiled one."""
    if (L, n_lat, sweeps_per_launch, threads) == (32, 1280, 2000, 256):
        name = "lds_sweep_kernel_L32_batched.md"
    elif (L, n_lat) == (256, 1280):
        name = "edge_slot_sweep_L256_batched.md"
    else:
        return {}
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    out = {}
    for line in open(path):
        key, sep, val = line.partition(":")
        if sep and key.isidentifier() and val.strip().isdigit():
            out[key] = int(val)
    return out


def cpu_baseline(specs, site_updates_per_thread, per_thread):
    """Oracle C restatement, one lattice per thread on every host core; each
    thread runs ``per_thread`` of the workload's lattices for the same number of
    sweeps, sized to about ``site_updates_per_thread`` site-updates."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import sw_oracle
    from rbim.clusters import activation_threshold

    sw_oracle.load_c()
    cores = len(os.sched_getaffinity(0))
    n = cores * per_thread
    chosen = [specs[(i * 37) % len(specs)] for i in range(n)]
    L = chosen[0].L
    sweeps = max(22, int(site_updates_per_thread) // (per_thread * L * L))

    def job(spec):
        s0 = sw_oracle.c_init_spins(spec.L, spec.seed, spec.lattice_id)
        sw_oracle.c_run(spec.L, s0, spec.J_h, spec.J_v, activation_threshold(spec.beta), spec.seed,
                        spec.lattice_id, 0, sweeps, sweeps // 11, 1, sweeps - sweeps // 11)
        return spec.L * spec.L * sweeps

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as pool:
        total = sum(pool.map(job, chosen))
    wall = time.perf_counter() - t0
    return {
        "value": total / wall,
        "unit": "site-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n} of the workload's lattices (L={L}) x {sweeps} sweeps, "
                  f"one lattice per thread, {wall:.1f} s wall",
    }


def fused_tile_size(L):
    """Tile side of the edge-slot sweep for L (0: L uses the multi-kernel tiled
    sweep); mirrors fused_tile_size() of the library."""
    return 64 if L >= 64 and L % 64 == 0 else 32 if L >= 32 and L % 32 == 0 else 0


def measure_point(L, args, steps, warmup, plain_steps, with_cpu):
    """Batched scan at one L (p, 40 T x 32 R = 1280 lattices by default), the
    plain one-lattice path and the CPU port beside it."""
    temperatures = np.linspace(1.0, 3.0, args.temps)
    specs = scan_specs(L, args.p, temperatures, args.realizations, seed=0)
    res = gpu_batched(specs, steps, warmup, args.engine, args.sweeps_per_launch, args.threads,
                      args.streams, args.tile)
    value = res["site_updates"] / (res["ms"] / 1e3)
    n_lat = len(specs)
    sites = L * L * n_lat
    lds = L <= 128 and args.engine != 1
    # every launch covers all lattices: one LDS kernel launch runs sweeps_per_launch
    # sweeps; the tiled path runs 4 kernels per sweep (timed together per sweep)
    launches = max(1, res["launches"])
    if lds:
        avg_launch_ms = res["ms"] / launches
        units = sites * steps / launches
        kernel = "lds_sweep_kernel"
    else:
        avg_launch_ms = res["ms"] / steps
        units = sites
        kernel = ("edge-slot sweep: ka_tile + kb_merge" if fused_tile_size(L) and args.engine != 3
                  else "tiled sweep: k1_tile_full + k2_merge + k3_flip + k5_finalize")
    achieved = ALGO_BYTES_PER_SITE_UPDATE * units / (avg_launch_ms / 1e3) / 1e9
    prof = profiled_fields(L, n_lat, res["sweeps_per_launch"], res["threads"])
    traffic = prof.get("traffic_bytes_per_launch" if lds else "traffic_bytes_per_sweep")
    valu = None
    if "valu_per_wavefront_sweep" in prof and "wavefronts_per_sweep" in prof:
        # what the kernel is bound by: VALU issue, from the profiled instruction
        # count over this run's time per sweep
        cycles = 4.0 * prof["valu_per_wavefront_sweep"] * prof["wavefronts_per_sweep"]
        valu = {"kernel": "lds_sweep_kernel" if lds else "ka_tile",
                "instructions_per_wavefront_sweep": prof["valu_per_wavefront_sweep"],
                "wavefronts_per_sweep": prof["wavefronts_per_sweep"],
                "frac": cycles / (VALU_SIMD_CYCLES_PER_S * res["ms"] / steps / 1e3)}
    plain = gpu_plain(specs[len(specs) // 2], plain_steps)
    point = {
        "L": L,
        "value": value,
        "ms_per_step": res["ms"] / steps,
        "steps": steps,
        "warmup": warmup,
        "workload": f"batched uniform scan: L={L}, p={args.p}, {args.temps} T in linspace(1,3) x "
                    f"{args.realizations} realizations = {n_lat} lattices, {steps} sweeps "
                    f"({run_schedule(steps)[0]} therm + {steps - run_schedule(steps)[0]} measured)",
        "lattices": n_lat,
        "threads_per_lattice": res["threads"],
        "sweeps_per_launch": res["sweeps_per_launch"],
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "kernel": kernel,
            "avg_launch_ms": avg_launch_ms,
            "launches": launches if lds else steps,
            "valu_issue": valu,
        },
        "plain": {"value": plain, "unit": "site-updates/s", "L": L,
                  "sample": f"one lattice, {plain_steps} sweeps, tiled kernels in hipGraphs"},
        "batched_vs_plain": value / plain,
    }
    if with_cpu:
        cpu = cpu_baseline(specs, args.cpu_site_updates, args.cpu_per_thread)
        point["cpu_baseline"] = cpu
        point["batched_vs_cpu"] = value / cpu["value"]
    return point


def workload_specs(name, seed=0):
    """Lattices of the ragged and large-lattice scans (one Batch each):
    ragged  L in {16, 32, 64, 128} x p in {0, 0.05, 0.10, 0.12} x 16 realizations x
            24 T in linspace(1.4, 2.8)  -> 6144 lattices, one stream per L;
    large   L in {512, 1024}, p = 0.10, 24 T in linspace(1.6, 2.6) x 8 realizations."""
    from rbim.engine import LatticeSpec

    specs = []
    if name == "ragged":
        for L in (16, 32, 64, 128):
            for i, p in enumerate((0.0, 0.05, 0.10, 0.12)):
                specs += scan_specs(L, p, np.linspace(1.4, 2.8, 24), 16, seed=seed + 10 * L + i)
    else:
        for L in (512, 1024):
            specs += scan_specs(L, 0.10, np.linspace(1.6, 2.6, 24), 8, seed=seed + L)
    # distinct lattice ids across the sub-scans
    return [LatticeSpec(L=s.L, J_h=s.J_h, J_v=s.J_v, beta=s.beta, seed=s.seed, lattice_id=k)
            for k, s in enumerate(specs)]


def measure_workload(name, steps):
    specs = workload_specs(name)
    res = gpu_batched(specs, steps, 0, 0)
    value = res["site_updates"] / (res["ms"] / 1e3)
    return {"workload": name, "lattices": len(specs), "steps": steps, "value": value,
            "unit": "site-updates/s", "ms_per_step": res["ms"] / steps,
            "Ls": sorted({s.L for s in specs})}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=22000, help="timed sweeps of every lattice")
    ap.add_argument("--warmup", type=int, default=200, help="untimed sweeps on an identical batch")
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--p", type=float, default=0.05)
    ap.add_argument("--temps", type=int, default=40)
    ap.add_argument("--realizations", type=int, default=32)
    ap.add_argument("--engine", type=int, default=0,
                    help="0 auto, 1 plain (tiled), 2 batched (LDS), 3 multi-kernel tiled")
    ap.add_argument("--sweeps-per-launch", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--streams", type=int, default=0, help="streams per tiled set (0: default)")
    ap.add_argument("--tile", type=int, default=0, help="tile side of the tiled engines (0: by L)")
    ap.add_argument("--plain-steps", type=int, default=400)
    ap.add_argument("--cpu-site-updates", type=float, default=2.5e8,
                    help="site-updates per host thread in the CPU sample (about 10 s)")
    ap.add_argument("--cpu-per-thread", type=int, default=4)
    ap.add_argument("--second-L", type=int, default=256,
                    help="second metric point, same recipe (0 to skip)")
    ap.add_argument("--second-steps", type=int, default=2200)
    ap.add_argument("--second-warmup", type=int, default=20)
    ap.add_argument("--second-plain-steps", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["ragged", "large"], default=None,
                    help="only time the ragged (L 16..128 mix) or large (L 512, 1024) scan")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (rocprof counter "
                         "calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_dispatch": 4 * n, "dispatches": 3}))
        return
    if args.workload:
        print(json.dumps(measure_workload(args.workload, args.steps)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): the lattices are "
              f"independent and shard across GPUs with no communication, so run one bench "
              f"per GPU; an N-GPU figure is N x the 1-GPU one.", file=sys.stderr)
        sys.exit(2)

    main_pt = measure_point(args.L, args, args.steps, args.warmup, args.plain_steps, not args.no_cpu)
    line = {
        "metric": "SW site-updates/s on 1 MI355X, batched vs plain, at L=32 & L=256 (bit-exact)",
        "value": main_pt["value"],
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_pt["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8/int32 (FP64 moments)",
        "data": "synthetic: seeded +-J disorder (numpy, as the reference draws it), hot start",
        "config": {
            "workload": main_pt["workload"],
            "L": args.L,
            "lattices": main_pt["lattices"],
            "engine": {0: "auto", 1: "plain", 2: "batched", 3: "tiled"}[args.engine],
            "threads_per_lattice": main_pt["threads_per_lattice"],
            "sweeps_per_launch": main_pt["sweeps_per_launch"],
        },
        "roofline": main_pt["roofline"],
        "plain": main_pt["plain"],
        "batched_vs_plain": main_pt["batched_vs_plain"],
    }
    if "cpu_baseline" in main_pt:
        line["cpu_baseline"] = main_pt["cpu_baseline"]
        line["batched_vs_cpu"] = main_pt["batched_vs_cpu"]
    if args.second_L and args.second_L != args.L:
        # the metric's second point (multi-tile union-find in both modes)
        line[f"L{args.second_L}"] = measure_point(args.second_L, args, args.second_steps,
                                                  args.second_warmup, args.second_plain_steps,
                                                  not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
