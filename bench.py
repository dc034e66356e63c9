#!/usr/bin/env python3
"""Swendsen-Wang throughput of the MI355X engine (site-updates/s).

Headline workload (the JSON line's `value`): the batched uniform scan --
L=32, p=0.05, 40 temperatures in linspace(1.0, 3.0) x 32 disorder
realizations = 1280 lattices, one workgroup per lattice with everything in
LDS, moments accumulated on the device.  A step is one sweep of every lattice
of the batch; --steps sweeps are timed (2000 thermalisation + 20000 measured
by default).  Disorder is synthetic (seeded numpy draws, as the reference
generates it), spins start hot.

Besides the contract keys the line carries, each with its own `roofline`:
  plain         the one-lattice-at-a-time tiled path at the same L;
  cpu_baseline  the oracle's C restatement of the reference sweep on all host
                cores (one lattice per thread), on a bounded sample (~10 s);
  L256          the metric's second point, the same recipe at L=256 (1280
                lattices of 65536 sites, 200 + 2000 sweeps);
  cfg2_full     the headline scan's whole run (2000 + 20000 sweeps) whatever
                --steps is;
  cfg1          the plain-method scan end to end (L=16, p=0.05, 10 T x 8
                realizations, 400 + 1600 sweeps): one lattice at a time on the
                plain path, all 80 in one batch, and the CPU port, each with
                the host estimators;
  cfg3          the ragged scan (L in {16, 32, 64, 128} x p in {0, .05, .10,
                .12} x 24 T x 16 realizations = 6144 lattices, 1000 + 10000
                sweeps);
  cfg4          the large-lattice scan (L = 512 and 1024, p = 0.10, 24 T x 8
                realizations, 500 + 5000 sweeps) and the p = 0, L = 1024
                energies at T = 2.0, 2.269, 2.5 against Kaufman's exact values;
  cfg5          the serving stream (Poisson 20 req/s, L in {32..256}) at a
                stated sweep scale: p50 / p99 latency, site-updates/s, and the
                highest sustainable request rate from a saturated stream;
  cpu_anchor    single-core sweep time of the CPU port at L=32 and L=64 against
                the reference's published numba figures.
Each `roofline` gives the algorithmic HBM rate (4 B per site-update: spin
read + write, J_h + J_v read) of the section's dominant kernel against
MI355X's 8 TB/s, the measured HBM traffic per launch where a committed
rocprofv3 summary covers the configuration, and `bound`, the limiter that
binds: "valu" with `valu_busy` = the kernel's measured VALUBusy share
(rocprofv3, 100 * SQ_ACTIVE_INST_VALU / CU_NUM / GRBM_GUI_ACTIVE) where that
summary holds it, "launch" for the one-lattice path.

    python bench.py [--gpus 1] [--steps 22000] [--warmup 200] [--only main L256 cfg1 ...]
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
ALGO_BYTES_PER_SITE_UPDATE = 4
# Exact energy per site of the 1024 x 1024 periodic Ising ferromagnet
# (Kaufman's finite-lattice partition function, Ferdinand and Fisher 1969)
KAUFMAN_L1024 = {2.0: -1.7455646, 2.269: -1.4154795, 2.5: -1.1060792}
# The reference's published numba sweep times on one core (README.md:110-111)
README_SWEEP_MS = {32: 0.133, 64: 0.424}
SECTIONS = ("main", "L256", "cut", "cfg2_full", "cfg1", "cfg3", "cfg4", "cfg5", "cpu_anchor")


# ------------------------------------------------------------------ inputs
def scan_specs(L, p, temperatures, R, seed):
    from rbim.engine import engine_seed
    from rbim.disorder import generate_disorder
    from rbim.simulation import scan_lattices

    disorder_seeds, run_seeds = np.random.SeedSequence(seed).spawn(2)
    rng = np.random.default_rng(disorder_seeds)
    realizations = [generate_disorder(L, p, rng) for _ in range(R)]
    return scan_lattices(L, realizations, temperatures, engine_seed(run_seeds))


def workload_specs(name, seed=0):
    """Lattices of the ragged and large-lattice scans (one Batch each):
    ragged  L in {16, 32, 64, 128} x p in {0, 0.05, 0.10, 0.12} x 16 realizations x
            24 T in linspace(1.4, 2.8)  -> 6144 lattices;
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


def run_schedule(K, therm_fraction=11):
    """(n_therm, n_bins, sweeps_per_bin) with n_therm + n_bins * spb == K,
    about 1/therm_fraction thermalisation and up to 10 bins."""
    measured = max(1, K - K // therm_fraction)
    n_bins = min(10, measured)
    spb = measured // n_bins
    return K - n_bins * spb, n_bins, spb


# ------------------------------------------------------------- GPU timing
def gpu_batched(specs, K, W, engine=0, sweeps_per_launch=0, threads=0, streams=0, tile=0,
                therm_fraction=11, single_launch=False):
    from rbim.engine import Batch

    n_therm, n_bins, spb = run_schedule(K, therm_fraction)
    kw = dict(engine=engine, sweeps_per_launch=sweeps_per_launch, threads=threads,
              streams=streams, tile=tile, single_launch=single_launch)
    if W > 0:
        wn, wb, ws = run_schedule(W)
        with Batch(specs, n_therm=wn, n_bins=wb, sweeps_per_bin=ws, **kw) as warm:
            warm.run_timed(W)
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, **kw) as batch:
        t0 = time.perf_counter()
        ms = batch.run_timed(K)
        wall = time.perf_counter() - t0
        info = batch.info()
        acc = batch.moments()
    sites = sum(s.L * s.L for s in specs)
    return {"ms": ms, "wall_s": wall, "site_updates": sites * K, "sites": sites, "acc": acc,
            "launches": int(info.kernel_launches), "threads": int(info.threads),
            "sweeps_per_launch": int(info.sweeps_per_launch), "groups": int(info.n_groups),
            "checksum": int(np.bitwise_xor.reduce(acc.reshape(-1) * np.uint64(0x9E3779B97F4A7C15)))}


def plain_matches(spec, K, acc):
    """One lattice of a batched run rerun alone on the plain path (same sweeps,
    schedule and key): are its bin sums (all 20 slots) identical?"""
    from rbim.engine import Batch

    n_therm, n_bins, spb = run_schedule(K)
    with Batch([spec], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=1) as batch:
        batch.run(K)
        return bool(np.array_equal(batch.moments()[0], acc))


def gpu_plain(spec, K):
    from rbim.engine import Batch

    n_therm, n_bins, spb = run_schedule(K)
    with Batch([spec], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=1) as warm:
        warm.run_timed(min(K, 64))
    with Batch([spec], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=1) as batch:
        ms = batch.run_timed(K)
    return spec.L * spec.L * K / (ms / 1e3)


# -------------------------------------------------------------- roofline
def profiled_fields(name):
    """The `key: number` lines of a committed rocprofv3 summary (profiles/)."""
    path = os.path.join(ROOT, "profiles", name)
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        key, sep, val = line.partition(":")
        if not (sep and key.isidentifier()):
            continue
        try:
            out[key] = float(val) if "." in val else int(val)
        except ValueError:
            pass
    return out


def roofline(kernel, site_updates_per_launch, avg_launch_ms, launches, profile=None,
             traffic_key=None, bound=None):
    """Algorithmic HBM rate of one launch of the dominant kernel vs 8 TB/s, with
    the measured traffic and VALU busy share of a committed profile."""
    prof = profiled_fields(profile) if profile else {}
    achieved = ALGO_BYTES_PER_SITE_UPDATE * site_updates_per_launch / (avg_launch_ms / 1e3) / 1e9
    valu = prof.get("valu_busy_pct")
    return {
        "bound": bound or ("valu" if valu is not None else "hbm"),
        "achieved": achieved,
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": achieved / PEAK_HBM_GBS,
        "traffic": prof.get(traffic_key) if traffic_key else None,
        "kernel": kernel,
        "avg_launch_ms": avg_launch_ms,
        "launches": launches,
        "valu_busy": None if valu is None else valu / 100.0,
        "profile": profile if prof else None,
    }


# ------------------------------------------------------------ CPU baseline
def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import sw_oracle

    sw_oracle.load_c()
    return sw_oracle


def cpu_run(specs, sweeps, n_therm, threads):
    """The oracle's C restatement on ``threads`` host threads, one lattice per
    thread; returns (site-updates, wall seconds, moment sums per lattice)."""
    from concurrent.futures import ThreadPoolExecutor

    from rbim.clusters import activation_threshold

    sw = _oracle()
    measured = sweeps - n_therm
    n_bins = min(10, max(1, measured))
    spb = measured // n_bins

    def job(spec):
        s0 = sw.c_init_spins(spec.L, spec.seed, spec.lattice_id)
        _, acc = sw.c_run(spec.L, s0, spec.J_h, spec.J_v, activation_threshold(spec.beta),
                          spec.seed, spec.lattice_id, 0, n_therm + n_bins * spb, n_therm, spb,
                          n_bins)
        return acc

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as pool:
        accs = list(pool.map(job, specs))
    wall = time.perf_counter() - t0
    return sum(s.L * s.L for s in specs) * (n_therm + n_bins * spb), wall, accs


def cpu_baseline(specs, site_updates_per_thread, per_thread):
    """Oracle C restatement, one lattice per thread on every host core; each
    thread runs ``per_thread`` of the workload's lattices for the same number of
    sweeps, sized to about ``site_updates_per_thread`` site-updates."""
    cores = len(os.sched_getaffinity(0))
    n = cores * per_thread
    chosen = [specs[(i * 37) % len(specs)] for i in range(n)]
    L = chosen[0].L
    sweeps = max(22, int(site_updates_per_thread) // (per_thread * L * L))
    total, wall, _ = cpu_run(chosen, sweeps, sweeps // 11, cores)
    return {"value": total / wall, "unit": "site-updates/s", "cores": cores, "kind": "port",
            "sample": f"{n} of the workload's lattices (L={L}) x {sweeps} sweeps, "
                      f"one lattice per thread, {wall:.1f} s wall"}


# ---------------------------------------------------------------- sections
def uses_edge_slot(L):
    """Whether the default tiled engine runs L on the edge-slot sweep: every
    L >= 3, on tiles of 32 or 64 with the last row and column cut when the
    side does not divide L (below 32 the lattice is one cut tile of 32)."""
    return L >= 3


def measure_point(L, args, steps, warmup, plain_steps, with_cpu):
    """Batched scan at one L (p, 40 T x 32 R = 1280 lattices by default), the
    plain one-lattice path and the CPU port beside it."""
    temperatures = np.linspace(1.0, 3.0, args.temps)
    specs = scan_specs(L, args.p, temperatures, args.realizations, seed=0)
    res = gpu_batched(specs, steps, warmup, args.engine, args.sweeps_per_launch, args.threads,
                      args.streams, args.tile)
    value = res["site_updates"] / (res["ms"] / 1e3)
    n_lat = len(specs)
    from rbim.engine import auto_uses_lds

    lds = (args.engine == 2 and L <= 128) or (args.engine == 0 and auto_uses_lds(L, len(specs)))
    launches = max(1, res["launches"])
    default_shape = (args.engine, args.sweeps_per_launch, args.threads, args.tile) == (0, 0, 0, 0)
    if lds:
        # every LDS launch covers all lattices for up to sweeps_per_launch sweeps
        profile = "lds_sweep_kernel_L32_batched.md" if (L, n_lat) == (32, 1280) and default_shape else None
        roof = roofline("lds_sweep_kernel", res["sites"] * steps / launches, res["ms"] / launches,
                        launches, profile, "traffic_bytes_per_launch")
    else:
        # the tiled path's launch: one sweep of its kernels over all lattices
        profile = "edge_slot_sweep_L256_batched.md" if (L, n_lat) == (256, 1280) and default_shape else None
        kernel = ("edge-slot sweep: ka_tile + kb_merge" if uses_edge_slot(L) and args.engine != 3
                  else "tiled sweep: k1_tile_full + k2_merge + k3_flip + k5_finalize")
        roof = roofline(kernel, res["sites"], res["ms"] / steps, steps, profile,
                        "traffic_bytes_per_sweep")
    plain = gpu_plain(specs[len(specs) // 2], plain_steps)
    # bit-exactness of the headline pair: the middle lattice's whole batched run
    # rerun alone on the plain path
    same = (None if args.no_plain_check else
            plain_matches(specs[len(specs) // 2], steps, res["acc"][len(specs) // 2]))
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
        "roofline": roof,
        "plain": {"value": plain, "unit": "site-updates/s", "L": L,
                  "sample": f"one lattice, {plain_steps} sweeps, tiled kernels in hipGraphs"},
        "batched_vs_plain": value / plain,
        "plain_moments_identical": same,
    }
    if with_cpu:
        cpu = cpu_baseline(specs, args.cpu_site_updates, args.cpu_per_thread)
        point["cpu_baseline"] = cpu
        point["batched_vs_cpu"] = value / cpu["value"]
    return point


def measure_cut(args):
    """Batched scans at sides that neither 64 nor 32 divides: the edge-slot sweep on
    tiles whose last row and column are cut (1280 lattices, 100 + 1000 sweeps)."""
    out = {"workload": "batched uniform scans at L = 129, 200, 333: p=0.05, 40 T in "
                       "linspace(1,3) x 32 realizations = 1280 lattices each, 100 + 1000 sweeps",
           "unit": "site-updates/s", "points": {}}
    for L in (129, 200, 333):
        specs = scan_specs(L, args.p, np.linspace(1.0, 3.0, 40), 32, seed=0)
        gpu_batched(specs, 20, 0)
        res = gpu_batched(specs, 1100, 0)
        point = {"value": res["site_updates"] / (res["ms"] / 1e3), "ms_per_step": res["ms"] / 1100,
                 "groups": res["groups"]}
        if L == 200:
            point["roofline"] = roofline("edge-slot sweep on cut tiles: ka_tile (whole tiles) + "
                                         "ka_tile_cut + kb_merge", res["sites"], res["ms"] / 1100,
                                         1100, "edge_slot_sweep_L200_cut_tiles.md",
                                         "traffic_bytes_per_sweep", bound="valu")
            point["roofline"]["bound_basis"] = (
                "the tile pass's measured VALUBusy on whole tiles (101 % at L=256, "
                "edge_slot_sweep_L256_batched.md); this item's own counters are not recorded")
        out["points"][str(L)] = point
    return out


def measure_cfg2_full(args):
    """The headline scan's whole run: 2000 thermalisation + 20000 measured sweeps."""
    specs = scan_specs(32, 0.05, np.linspace(1.0, 3.0, 40), 32, seed=0)
    res = gpu_batched(specs, 22000, 0)
    launches = max(1, res["launches"])
    return {"workload": "L=32, p=0.05, 40 T x 32 R = 1280 lattices, 2000 + 20000 sweeps",
            "value": res["site_updates"] / (res["ms"] / 1e3), "unit": "site-updates/s",
            "wall_s": res["ms"] / 1e3,
            "roofline": roofline("lds_sweep_kernel", res["sites"] * 22000 / launches,
                                 res["ms"] / launches, launches, "lds_sweep_kernel_L32_batched.md",
                                 "traffic_bytes_per_launch")}


def _scan_results(L, temps, R, accs, n_bins):
    """Host estimators over per-lattice moment sums (task order)."""
    from rbim.engine import moments_to_primitive_arrays
    from rbim.simulation import _assemble_arrays

    prims, bins = moments_to_primitive_arrays(np.stack(accs))
    return _assemble_arrays(L, temps, prims, bins, R, {})


def measure_cfg1(args):
    """Plain method, end to end: L=16, p=0.05, 10 T x 8 R, 400 + 1600 sweeps,
    moments to the results dict (host estimators included)."""
    from rbim.engine import Batch

    L, temps, R, n_therm, n_sweeps, n_bins = 16, np.linspace(1.0, 3.0, 10), 8, 400, 1600, 10
    spb = n_sweeps // n_bins
    specs = scan_specs(L, 0.05, temps, R, seed=0)
    total = n_therm + n_bins * spb
    site_updates = len(specs) * L * L * total
    # warm the plain path's kernels and graphs
    with Batch([specs[0]], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=1) as warm:
        warm.run_timed(64)
    t0 = time.perf_counter()
    accs, kernel_ms = [], 0.0
    for spec in specs:  # one lattice at a time, as the reference's plain method
        with Batch([spec], n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=1) as b:
            kernel_ms += b.run_timed(total)
            accs.append(b.moments()[0])
    res_plain = _scan_results(L, temps, R, accs, n_bins)
    plain_wall = time.perf_counter() - t0
    t0 = time.perf_counter()
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb) as b:
        batched_ms = b.run_timed(total)
        acc_b = b.moments()
    res_batched = _scan_results(L, temps, R, list(acc_b), n_bins)
    batched_wall = time.perf_counter() - t0
    identical = bool(np.array_equal(np.stack(accs), acc_b))
    out = {
        "workload": "plain scan: L=16, p=0.05, 10 T in linspace(1,3) x 8 realizations = 80 "
                    "lattices, 400 therm + 1600 measured sweeps, 10 bins, hot start",
        "site_updates": site_updates,
        "plain_wall_s": plain_wall,
        "plain_value": site_updates / plain_wall,
        "batched_wall_s": batched_wall,
        "batched_value": site_updates / batched_wall,
        "unit": "site-updates/s",
        "plain_and_batched_moments_identical": identical,
        "energy_T1": float(res_plain["observables"]["energy"]["value"][0]),
        # a launch here: one sweep of the plain path's kernels on one lattice
        "roofline": roofline("edge-slot sweep (one lattice, one cut tile of 32): ka_tile_cut + "
                             "kb_merge",
                             L * L, kernel_ms / (len(specs) * total), len(specs) * total,
                             bound="launch"),
        "reference_readme_s": 5.2,
        "reference_readme_note": "README.md:112: n_sweeps=2000 with the default 200 therm, "
                                 "process pool of os.cpu_count() workers, hardware unstated",
    }
    out["energy_T1_batched"] = float(res_batched["observables"]["energy"]["value"][0])
    if not args.no_cpu:
        cores = len(os.sched_getaffinity(0))
        cpu_total, cpu_wall, cpu_accs = cpu_run(specs, total, n_therm, cores)
        out["cpu_baseline"] = {"value": cpu_total / cpu_wall, "unit": "site-updates/s",
                               "cores": cores, "kind": "port", "wall_s": cpu_wall,
                               "sample": "the whole scan (80 lattices), one lattice per thread"}
        out["cpu_moments_identical"] = bool(np.array_equal(np.stack(cpu_accs), acc_b))
    return out


def measure_cfg3(args):
    """The ragged scan as one launch sequence per L (each on its own stream, each
    side on the auto engine's choice) and as one launch over every L on the LDS
    kernel (lds_ragged_kernel: a page of lattices of one L per 1024-thread
    workgroup)."""
    specs = workload_specs("ragged")
    out = {
        "workload": "ragged scan: L in {16,32,64,128} x p in {0,0.05,0.10,0.12} x 24 T in "
                    "linspace(1.4,2.8) x 16 realizations = 6144 lattices, 1000 + 10000 sweeps",
        "lattices": len(specs), "unit": "site-updates/s",
    }
    by_L = {L: sum(s.L * s.L for s in specs if s.L == L) for L in (16, 32, 64, 128)}
    for single in (False, True):
        # the single launch holds every side on the LDS kernel (L = 128 too),
        # the default picks each side's engine (the edge-slot sweep at L = 128)
        engine = 2 if single else 0
        gpu_batched(specs, 20, 0, engine=engine, single_launch=single)  # load the kernels
        res = gpu_batched(specs, 11000, 0, engine=engine, therm_fraction=11, single_launch=single)
        launches = max(1, res["launches"])
        sites = res["sites"]
        item = {
            "value": res["site_updates"] / (res["ms"] / 1e3), "wall_s": res["ms"] / 1e3,
            "groups": res["groups"], "launches": launches, "checksum": res["checksum"],
            "roofline": (roofline("lds_ragged_kernel (every L in one launch)",
                                  sites * 11000 / launches, res["ms"] / launches, launches,
                                  "lds_ragged_kernel_ragged_scan.md")
                         if single else
                         roofline("one launch sequence per L, concurrent: edge-slot sweep "
                                  "(ka_tile + kb_merge) for L = 128, lds_sweep_kernel below",
                                  sites * 11000, res["ms"], 1,
                                  "edge_slot_sweep_and_lds_sweep_kernel_ragged_scan.md")),
        }
        out["single_launch" if single else "per_L_streams"] = item
    out["site_share"] = {str(L): by_L[L] / sum(by_L.values()) for L in by_L}
    out["moments_identical"] = out["single_launch"]["checksum"] == out["per_L_streams"]["checksum"]
    best = max(("per_L_streams", "single_launch"), key=lambda k: out[k]["value"])
    out["value"] = out[best]["value"]
    out["default"] = "per_L_streams"
    return out


def measure_cfg4(args):
    from rbim.simulation import run_temperature_scan

    specs = workload_specs("large")
    gpu_batched(specs, 8, 0)
    res = gpu_batched(specs, 5500, 0)
    out = {
        "workload": "large scan: L in {512, 1024}, p=0.10, 24 T in linspace(1.6,2.6) x 8 "
                    "realizations = 384 lattices, 500 + 5000 sweeps",
        "lattices": len(specs), "value": res["site_updates"] / (res["ms"] / 1e3),
        "unit": "site-updates/s", "wall_s": res["ms"] / 1e3,
        "roofline": roofline("edge-slot sweep, slot merge in HBM: ka_tile + kg_union + kg_fold + "
                             "kg_flip + kg_sweep_end", res["sites"], res["ms"] / 5500, 5500,
                             "edge_slot_sweep_L512_L1024.md", "traffic_bytes_per_sweep"),
    }
    temps = sorted(KAUFMAN_L1024)
    t0 = time.perf_counter()
    scan = run_temperature_scan(1024, 0.0, temps, n_realizations=4, n_sweeps=1000, n_therm=200,
                                n_bins=10, seed=3, start="cold")
    e = scan["observables"]["energy"]
    rows = []
    for t, T in enumerate(temps):
        err = float(max(e["error"][t], e["thermal_error"][t]))
        rows.append({"T": T, "energy": float(e["value"][t]), "error": err,
                     "kaufman": KAUFMAN_L1024[T],
                     "z": float((e["value"][t] - KAUFMAN_L1024[T]) / err) if err > 0 else None})
    out["onsager_p0_L1024"] = {"sample": "4 replicas x 3 T, cold start, 200 + 1000 sweeps",
                               "wall_s": time.perf_counter() - t0, "points": rows,
                               "max_abs_z": max(abs(r["z"]) for r in rows if r["z"] is not None)}
    return out


def measure_cfg5(args):
    from rbim.serving import ScanServer, mean_request_site_updates, poisson_requests, replay

    out = {"stream": "Poisson 20 req/s, L in {32,64,128,256}, p ~ U[0,0.12], 8-40 T in [1,3], "
                     "4-32 realizations, 1000-20000 sweeps x sweep_scale"}
    server = ScanServer(max_active=16, chunk_sweeps=500)
    warm = poisson_requests(20.0, 0.3, seed=99)
    replay(server, warm, 0.01)
    server.close()
    # latency at a load the GPU absorbs
    server = ScanServer(max_active=16, chunk_sweeps=500)
    arrivals = poisson_requests(20.0, args.serve_duration, seed=0)
    stats, offered, wall = replay(server, arrivals, args.serve_scale)
    server.close()
    out["latency_run"] = {"sweep_scale": args.serve_scale, "duration_s": args.serve_duration,
                          "requests": len(arrivals), "offered_site_updates": float(offered),
                          "offered_rate": float(offered) / args.serve_duration,
                          "p50_latency_s": stats["p50_latency_s"],
                          "p99_latency_s": stats["p99_latency_s"],
                          "site_updates_per_s": stats["site_updates_per_s"], "wall_s": wall,
                          "peak_active": stats["peak_active"]}
    # capacity: a stream that offers several times what the GPU absorbs
    server = ScanServer(max_active=16, chunk_sweeps=500)
    arrivals = poisson_requests(20.0, args.saturate_duration, seed=1)
    stats2, offered2, wall2 = replay(server, arrivals, args.saturate_scale)
    server.close()
    capacity = stats2["site_updates_per_s"]
    per_request = mean_request_site_updates()
    out["saturated_run"] = {"sweep_scale": args.saturate_scale,
                            "duration_s": args.saturate_duration, "requests": len(arrivals),
                            "offered_site_updates": float(offered2),
                            "site_updates_per_s": capacity, "wall_s": wall2,
                            "p50_latency_s": stats2["p50_latency_s"],
                            "p99_latency_s": stats2["p99_latency_s"]}
    out["mean_request_site_updates_full_scale"] = per_request
    out["sustainable_rate_hz"] = {"full_scale": capacity / per_request,
                                  str(args.serve_scale): capacity / (per_request * args.serve_scale)}
    # the saturated run: what the kernels sustain (the latency run serves what
    # the Poisson stream offers, so its rate is set by the arrivals)
    out["roofline"] = roofline("serving mix, saturated (LDS kernel L <= 128, edge-slot sweep "
                               "L = 256)", stats2["site_updates"], wall2 * 1e3, 1,
                               bound="valu")
    out["roofline"]["bound_basis"] = (
        "the mix runs the LDS kernel and the edge-slot sweep, each VALU-bound in its own "
        "profile (lds_sweep_kernel_L32_batched.md, edge_slot_sweep_L256_batched.md); the "
        "latency run is bound by its arrivals")
    return out


def measure_cpu_anchor(args):
    """Single-core sweep time of the CPU port against README's numba figures."""
    out = {}
    for L in (32, 64):
        spec = scan_specs(L, 0.05, [2.0], 1, seed=0)[0]
        sweeps = int(4.0e7 // (L * L))
        total, wall, _ = cpu_run([spec], sweeps, sweeps // 10, 1)
        ms = wall * 1e3 / sweeps
        out[f"L{L}"] = {"ms_per_sweep": ms, "reference_numba_ms": README_SWEEP_MS[L],
                        "faster_than_reference": README_SWEEP_MS[L] / ms, "sweeps": sweeps,
                        "T": 2.0, "cores": 1}
    return out


def measure_workload(name, steps, single_launch=False):
    specs = workload_specs(name)
    # the single launch holds every side on the LDS kernel, as in cfg3
    res = gpu_batched(specs, steps, 0, engine=2 if single_launch else 0, single_launch=single_launch)
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
    ap.add_argument("--serve-scale", type=float, default=0.02)
    ap.add_argument("--serve-duration", type=float, default=4.0)
    ap.add_argument("--saturate-scale", type=float, default=0.2)
    ap.add_argument("--saturate-duration", type=float, default=1.5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain-check", action="store_true",
                    help="skip the plain-path rerun behind plain_moments_identical (profiling "
                         "runs: under rocprofv3 --kernel-trace its extra graphs ended in a "
                         "crash inside the HIP runtime)")
    ap.add_argument("--only", nargs="+", choices=SECTIONS, default=None,
                    help="sections to run (default: all)")
    ap.add_argument("--workload", choices=["ragged", "large"], default=None,
                    help="only time the ragged (L 16..128 mix) or large (L 512, 1024) scan")
    ap.add_argument("--single-launch", action="store_true",
                    help="with --workload ragged: every L in one launch")
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
        print(json.dumps(measure_workload(args.workload, args.steps, args.single_launch)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): the lattices are "
              f"independent and shard across GPUs with no communication, so run one bench "
              f"per GPU; an N-GPU figure is N x the 1-GPU one.", file=sys.stderr)
        sys.exit(2)
    only = set(args.only or SECTIONS)
    t_start = time.perf_counter()

    def note(section):
        print(f"bench: {section} done at {time.perf_counter() - t_start:.1f} s", file=sys.stderr,
              flush=True)

    line = {
        "metric": "SW site-updates/s on 1 MI355X, batched vs plain, at L=32 & L=256 (bit-exact)",
        "value": None,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8/int32 (FP64 moments)",
        "data": "synthetic: seeded +-J disorder (numpy, as the reference draws it), hot start",
    }
    if "main" in only:
        main_pt = measure_point(args.L, args, args.steps, args.warmup, args.plain_steps,
                                not args.no_cpu)
        line.update({
            "value": main_pt["value"],
            "ms_per_step": main_pt["ms_per_step"],
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
            "plain_moments_identical": main_pt["plain_moments_identical"],
        })
        if "cpu_baseline" in main_pt:
            line["cpu_baseline"] = main_pt["cpu_baseline"]
            line["batched_vs_cpu"] = main_pt["batched_vs_cpu"]
        note("main")
    if "L256" in only and args.second_L and args.second_L != args.L:
        # the metric's second point (multi-tile union-find in both modes)
        line[f"L{args.second_L}"] = measure_point(args.second_L, args, args.second_steps,
                                                  args.second_warmup, args.second_plain_steps,
                                                  not args.no_cpu)
        note("L256")
    if "cut" in only:
        line["cut"] = measure_cut(args)
        ref = line.get(f"L{args.second_L}", {}).get("value")
        if ref:
            for pt in line["cut"]["points"].values():
                pt[f"vs_L{args.second_L}"] = pt["value"] / ref
        note("cut")
    for key, fn in (("cfg2_full", measure_cfg2_full), ("cfg1", measure_cfg1),
                    ("cfg3", measure_cfg3), ("cfg4", measure_cfg4), ("cfg5", measure_cfg5)):
        if key in only:
            line[key] = fn(args)
            note(key)
    if "cpu_anchor" in only and not args.no_cpu:
        line["cpu_anchor"] = measure_cpu_anchor(args)
        note("cpu_anchor")
    line["bench_wall_s"] = time.perf_counter() - t_start
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
