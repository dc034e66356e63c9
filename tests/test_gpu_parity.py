"""Bit-exact parity of the HIP engine with the oracle, through the C ABI.

Both engines are checked: ENGINE_PLAIN (tiled kernels, any L) and
ENGINE_BATCHED (LDS-resident kernel, L <= 128).  Given identical spins,
couplings, occupation and flip bits, labels (reference numbering and min-site
canonical form), n_clusters, wrap_x / wrap_y, sizes, M_C, E and M must be
identical; full trajectories (spins and the 20-slot moment sums, including the
FP64 |M_C|/S_max sum) must be identical sweep for sweep.
"""

import ctypes

import numpy as np
import pytest

from conftest import random_disorder
from samples import column_case, diagonal_loop, double_winding, sample_configurations

pytestmark = pytest.mark.gpu

PLAIN, BATCHED = 1, 2


@pytest.fixture(scope="module")
def hip():
    from rbim import _hip

    _hip.context()
    return _hip


def gpu_label(hip, active, L, engine):
    act = np.ascontiguousarray(active, dtype=np.uint8)
    labels = np.empty(L * L, np.int32)
    canon = np.empty(L * L, np.int32)
    n = ctypes.c_int32()
    wraps = np.zeros(2, np.int32)
    ctx = hip.context()
    hip.check(hip.lib().rbim_label_and_wrap(ctx, L, hip.ptr(act), engine, hip.ptr(labels),
                                             hip.ptr(canon), ctypes.byref(n), hip.ptr(wraps)), ctx)
    return n.value, labels, canon, bool(wraps[0]), bool(wraps[1])


def gpu_sweep_given_bits(hip, s, engine):
    L = s["L"]
    spins = s["spins"].copy()
    act = np.ascontiguousarray(s["active"], np.uint8)
    flips = np.ascontiguousarray(s["flips"], np.uint8)
    labels = np.empty(L * L, np.int32)
    st = hip.SweepStats()
    ctx = hip.context()
    hip.check(hip.lib().rbim_sweep_given_bits(ctx, L, hip.ptr(spins), hip.ptr(s["J_h"]),
                                               hip.ptr(s["J_v"]), hip.ptr(act), hip.ptr(flips),
                                               engine, hip.ptr(labels), ctypes.byref(st)), ctx)
    stats = {name: getattr(st, name) for name, _ in hip.SweepStats._fields_}
    stats["wrap_x"] = bool(stats["wrap_x"])
    stats["wrap_y"] = bool(stats["wrap_y"])
    return spins, stats, labels


SIZES = (3, 4, 5, 6, 8, 10, 16, 24, 32, 64)


@pytest.mark.parametrize("engine", [PLAIN, BATCHED])
def test_label_and_wrap_bit_exact(hip, coracle, engine):
    checked = wrapped = 0
    for s in sample_configurations(coracle, SIZES, (0.0, 0.1, 0.5), (1.0, 2.27, 4.0), 12, seed=2024):
        L = s["L"]
        n_ref, lab_ref, wx_ref, wy_ref = coracle.c_label_and_wrap(s["active"], L)
        n, labels, canon, wx, wy = gpu_label(hip, s["active"], L, engine)
        case = f"L={L} p={s['p']} T={s['T']}"
        assert n == n_ref, case
        assert np.array_equal(labels, lab_ref), case
        assert np.array_equal(canon, coracle.canonical(lab_ref)), case
        assert (wx, wy) == (wx_ref, wy_ref), case
        checked += 1
        wrapped += int(wx_ref or wy_ref)
    assert checked >= 1000 and wrapped > 100


@pytest.mark.parametrize("engine", [PLAIN, BATCHED])
def test_label_known_answers(hip, engine):
    L = 8
    none = np.zeros(2 * L * L, bool)
    assert gpu_label(hip, none, L, engine)[0] == L * L
    assert gpu_label(hip, none, L, engine)[3:] == (False, False)
    n, _, canon, wx, wy = gpu_label(hip, ~none, L, engine)
    assert n == 1 and wx and wy and np.all(canon == 0)
    open_col, closed_col = column_case(6)
    assert gpu_label(hip, open_col, 6, engine)[3:] == (False, False)
    assert gpu_label(hip, closed_col, 6, engine)[3:] == (False, True)
    for L in (3, 4, 7, 40):
        assert gpu_label(hip, diagonal_loop(L), L, engine)[3:] == (True, True)
    assert gpu_label(hip, double_winding(6), 6, engine)[3:] == (False, True)


def test_label_and_wrap_large_lattices(hip, coracle):
    """Tiled path with several tiles per side, partial edge tiles and odd L."""
    checked = 0
    # tile sides: 32 with partial edge tiles (33, 47, 100, 129), 32 full (96),
    # 64 full (128, 192, 256)
    for s in sample_configurations(coracle, (33, 47, 96, 100, 128, 129, 192, 256), (0.0, 0.12),
                                   (1.5, 2.27, 3.0), 2, seed=77):
        L = s["L"]
        n_ref, lab_ref, wx_ref, wy_ref = coracle.c_label_and_wrap(s["active"], L)
        n, labels, canon, wx, wy = gpu_label(hip, s["active"], L, PLAIN)
        assert (n, wx, wy) == (n_ref, wx_ref, wy_ref), f"L={L}"
        assert np.array_equal(labels, lab_ref), f"L={L}"
        checked += 1
    assert checked == 96


def test_activation_bit_exact(hip, coracle):
    rng = np.random.default_rng(31)
    for L in (3, 7, 16, 33, 64, 130):
        J_h, J_v = random_disorder(rng, L, 0.2)
        spins = np.where(rng.random(L * L) < 0.5, -1, 1).astype(np.int8)
        for beta, key, lid, sweep in [(0.0, 1, 0, 0), (0.45, 2**63 + 5, 9, 77), (50.0, 3, 2, 5)]:
            thr = coracle.threshold31(beta)
            a_ref, f_ref = coracle.c_activate(spins, J_h, J_v, thr, key, lid, sweep)
            active = np.empty(2 * L * L, np.uint8)
            flips = np.empty(L * L, np.uint8)
            ctx = hip.context()
            hip.check(hip.lib().rbim_activate(ctx, L, hip.ptr(spins), hip.ptr(J_h), hip.ptr(J_v),
                                               thr, key, lid, sweep, hip.ptr(active),
                                               hip.ptr(flips)), ctx)
            assert np.array_equal(active.astype(bool), a_ref), f"L={L} beta={beta}"
            assert np.array_equal(flips, f_ref), f"L={L} beta={beta}"


@pytest.mark.parametrize("engine", [PLAIN, BATCHED])
def test_sweep_given_bits_bit_exact(hip, coracle, engine):
    n = 0
    for s in sample_configurations(coracle, (3, 5, 8, 16, 32, 48, 64), (0.0, 0.1), (1.2, 2.27, 3.5), 3,
                                   seed=99):
        sp_ref, st_ref, lab_ref = coracle.c_sweep_given_bits(s["spins"], s["J_h"], s["J_v"],
                                                              s["active"], s["flips"])
        sp, st, labels = gpu_sweep_given_bits(hip, s, engine)
        case = f"L={s['L']} p={s['p']} T={s['T']}"
        assert np.array_equal(sp, sp_ref), case
        assert np.array_equal(labels, lab_ref), case
        assert st == st_ref, case
        n += 1
    assert n == 126


def _oracle_trajectory(coracle, spec, n_sweeps, n_therm, spb, n_bins, start="hot"):
    from rbim.clusters import activation_threshold

    s0 = coracle.c_init_spins(spec.L, spec.seed, spec.lattice_id, start)
    return coracle.c_run(spec.L, s0, spec.J_h, spec.J_v, activation_threshold(spec.beta), spec.seed,
                         spec.lattice_id, 0, n_sweeps, n_therm, spb, n_bins)


def _specs(Ls, seed=0, p=0.1, temps=(1.4, 2.27, 3.2)):
    from rbim.engine import LatticeSpec

    rng = np.random.default_rng(seed)
    specs = []
    k = 0
    for L in Ls:
        J_h, J_v = random_disorder(rng, L, p)
        for T in temps:
            specs.append(LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / T,
                                     seed=0x243F6A8885A308D3 + 17 * L, lattice_id=k))
            k += 1
    return specs


@pytest.mark.parametrize("engine", [PLAIN, BATCHED])
def test_trajectories_bit_exact(hip, coracle, engine):
    from rbim.engine import Batch

    specs = _specs((3, 5, 16, 17, 24, 32, 48, 64))
    n_therm, spb, n_bins = 10, 9, 5
    total = n_therm + spb * n_bins
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=engine,
               sweeps_per_launch=7) as batch:
        batch.run(13)  # split runs: bins and graphs straddle calls
        batch.run(total - 13)
        acc = batch.moments()
        spins = batch.spins()
    for i, spec in enumerate(specs):
        s_ref, acc_ref = _oracle_trajectory(coracle, spec, total, n_therm, spb, n_bins)
        assert np.array_equal(spins[i], s_ref), f"lattice {i} L={spec.L}"
        assert np.array_equal(acc[i], acc_ref), f"lattice {i} L={spec.L}"


def test_lds_kernel_l128_and_cold_start(hip, coracle):
    from rbim.engine import Batch

    specs = _specs((100, 128), temps=(2.0, 2.6))
    n_therm, spb, n_bins = 4, 5, 2
    total = n_therm + spb * n_bins
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=BATCHED,
               start="cold") as batch:
        batch.run()
        acc = batch.moments()
        spins = batch.spins()
    for i, spec in enumerate(specs):
        s_ref, acc_ref = _oracle_trajectory(coracle, spec, total, n_therm, spb, n_bins, start="cold")
        assert np.array_equal(spins[i], s_ref) and np.array_equal(acc[i], acc_ref), f"L={spec.L}"


def test_tiled_large_lattices_bit_exact(hip, coracle):
    from rbim.engine import Batch

    specs = _specs((160, 256, 512), temps=(2.0, 2.4), seed=4)
    n_therm, spb, n_bins = 3, 3, 2
    total = n_therm + spb * n_bins
    with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=PLAIN,
               sweeps_per_launch=4) as batch:
        batch.run()
        acc = batch.moments()
        spins = batch.spins()
    for i, spec in enumerate(specs):
        s_ref, acc_ref = _oracle_trajectory(coracle, spec, total, n_therm, spb, n_bins)
        assert np.array_equal(spins[i], s_ref) and np.array_equal(acc[i], acc_ref), f"L={spec.L}"


def test_batched_equals_plain_and_is_launch_shape_invariant(hip):
    from rbim.engine import Batch

    specs = _specs((16, 32), seed=8, temps=(1.8, 2.3, 2.9))
    out = []
    for engine, threads, per_launch in [(PLAIN, 0, 5), (BATCHED, 64, 3), (BATCHED, 256, 50),
                                        (BATCHED, 1024, 1000)]:
        with Batch(specs, n_therm=20, n_bins=4, sweeps_per_bin=10, engine=engine, threads=threads,
                   sweeps_per_launch=per_launch) as batch:
            batch.run()
            out.append((batch.moments(), batch.spins()))
    for acc, spins in out[1:]:
        assert np.array_equal(acc, out[0][0])
        for a, b in zip(spins, out[0][1]):
            assert np.array_equal(a, b)
    # a lattice's result does not depend on the rest of its batch
    with Batch(specs[2:3], n_therm=20, n_bins=4, sweeps_per_bin=10, engine=BATCHED) as alone:
        alone.run()
        assert np.array_equal(alone.moments()[0], out[0][0][2])


@pytest.mark.parametrize("engine", [PLAIN, BATCHED, 3])
def test_random_shapes_bit_exact(hip, coracle, engine):
    """A seeded draw of shapes the fixed lists above do not hold: ten or so
    lattices with L from 3 to 140 (odd, even and multiples of 32 mixed in one
    batch; L <= 128 for the LDS engine), p up to 0.2, T in [1, 4], random seeds
    and lattice ids, hot and cold starts; trajectories (spins, all 20 moment
    slots) equal the oracle's."""
    from rbim.engine import Batch, LatticeSpec

    rng = np.random.default_rng(20261203 + engine)
    top = 128 if engine == BATCHED else 140
    Ls = sorted({int(x) for x in rng.integers(3, top + 1, size=9)} | {32, 96 if top > 96 else 64})
    specs = []
    for k, L in enumerate(Ls):
        J_h, J_v = random_disorder(rng, L, float(rng.uniform(0.0, 0.2)))
        specs.append(LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / float(rng.uniform(1.0, 4.0)),
                                 seed=int(rng.integers(0, 2**63)), lattice_id=int(rng.integers(0, 2**31))))
    n_therm, spb, n_bins = 4, 3, 3
    total = n_therm + spb * n_bins
    for start in ("hot", "cold"):
        with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=engine,
                   start=start) as batch:
            batch.run(5)
            batch.run(total - 5)
            acc = batch.moments()
            spins = batch.spins()
        for i, spec in enumerate(specs):
            s_ref, acc_ref = _oracle_trajectory(coracle, spec, total, n_therm, spb, n_bins, start=start)
            assert np.array_equal(spins[i], s_ref), (engine, start, spec.L)
            assert np.array_equal(acc[i], acc_ref), (engine, start, spec.L)


def test_single_launch_ragged_batch_bit_exact(hip, coracle):
    """One launch over L = 16, 32, 64, 128 (pages of 16, 4, 2 and 1 lattices
    per workgroup, last pages part-filled) beside a tiled lattice of L = 160:
    trajectories equal the oracle's and the one-launch-per-L batch's."""
    from rbim.engine import Batch, LatticeSpec

    rng = np.random.default_rng(2026)
    specs = []
    for L, count in ((16, 17), (32, 5), (64, 3), (128, 2), (160, 1)):
        for _ in range(count):
            J_h, J_v = random_disorder(rng, L, float(rng.uniform(0.0, 0.15)))
            specs.append(LatticeSpec(L=L, J_h=J_h, J_v=J_v, beta=1.0 / float(rng.uniform(1.2, 3.5)),
                                     seed=int(rng.integers(0, 2**63)),
                                     lattice_id=int(rng.integers(0, 2**31))))
    rng.shuffle(specs)
    n_therm, spb, n_bins = 3, 4, 2
    total = n_therm + spb * n_bins
    out = {}
    for single in (True, False):
        with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, single_launch=single,
                   sweeps_per_launch=4 if single else 0) as batch:
            batch.run(5)
            batch.run(total - 5)
            out[single] = (batch.moments(), batch.spins(), batch.info())
    acc, spins, info = out[True]
    assert info.lds_groups == 4 and info.n_groups == 5  # the one tiled lattice: one part
    assert np.array_equal(acc, out[False][0])
    for i, spec in enumerate(specs):
        s_ref, acc_ref = _oracle_trajectory(coracle, spec, total, n_therm, spb, n_bins)
        assert np.array_equal(spins[i], s_ref), spec.L
        assert np.array_equal(acc[i], acc_ref), spec.L
        assert np.array_equal(out[False][1][i], s_ref), spec.L


def test_lds_sets_split_over_streams_are_invariant(hip):
    """An LDS set split over streams -- explicitly, or by default when it fills
    the GPU more than once with a part-empty last round (1100 lattices at L=64,
    about 512 resident at a time) -- gives the moments and spins of the
    unsplit set: every lattice is keyed by its own id."""
    from rbim.engine import Batch, LatticeSpec

    def run(specs, streams):
        with Batch(specs, n_therm=3, n_bins=2, sweeps_per_bin=4, engine=BATCHED,
                   streams=streams) as batch:
            batch.run()
            return batch.moments(), batch.spins()

    small = _specs((16, 32), seed=3, temps=(1.8, 2.3, 2.9, 3.4))
    ref_acc, ref_sp = run(small, 1)
    acc, sp = run(small, 3)
    assert np.array_equal(acc, ref_acc) and all(np.array_equal(a, b) for a, b in zip(sp, ref_sp))
    rng = np.random.default_rng(4)
    J_h, J_v = random_disorder(rng, 64, 0.1)
    many = [LatticeSpec(L=64, J_h=J_h, J_v=J_v, beta=1.0 / (1.5 + 1.5 * k / 1100), seed=11,
                        lattice_id=k) for k in range(1100)]
    ref_acc, ref_sp = run(many, 1)
    acc, sp = run(many, 0)
    assert np.array_equal(acc, ref_acc) and all(np.array_equal(a, b) for a, b in zip(sp, ref_sp))


def test_tiled_sets_split_over_streams_are_invariant(hip):
    """Tiled sets of two sizes (edge-slot sweep, and the multi-kernel sweep)
    split over one, two (the default) or three streams: a run queues one
    launch per group in turn -- graphs of three sweeps and single sweeps mixed,
    runs ending inside a graph's length -- and every split gives the moments,
    spins and per-sweep series of the unsplit set."""
    from rbim._hip import ENGINE_TILED
    from rbim.engine import Batch

    specs = _specs((96, 160), seed=21, temps=(1.6, 2.3, 3.0))
    for engine in (PLAIN, ENGINE_TILED):
        out = []
        for streams in (1, 0, 3):
            with Batch(specs, n_therm=4, n_bins=2, sweeps_per_bin=5, engine=engine, streams=streams,
                       sweeps_per_launch=3, series=10) as batch:
                batch.run(4)
                batch.run()
                out.append((batch.moments(), batch.spins(), batch.series()))
        for acc, spins, series in out[1:]:
            assert np.array_equal(acc, out[0][0]), engine
            assert np.array_equal(series, out[0][2]), engine
            assert all(np.array_equal(a, b) for a, b in zip(spins, out[0][1])), engine


def test_run_past_end_is_refused(hip):
    from rbim.engine import Batch
    from rbim._hip import EngineError

    specs = _specs((8,), temps=(2.0,))
    with Batch(specs, n_therm=1, n_bins=1, sweeps_per_bin=2) as batch:
        batch.run(3)
        with pytest.raises(EngineError):
            batch.run(1)


@pytest.mark.parametrize("Ls", [(32, 64, 96), (128, 160, 256), (320,), (288, 384),
                                (33, 47, 100, 129), (200,), (333,), (226, 131), (352,),
                                (65, 127, 191), (193, 255), (257,), (130, 321),
                                (3, 4, 5, 8), (16, 17, 31)])
def test_edge_slot_sweep_equals_hbm_forest_sweep(hip, coracle, Ls):
    """The edge-slot sweep (ENGINE_PLAIN) on its default tiles and
    on tiles of 32 and of 64 forced, and the multi-kernel HBM-forest sweep
    (ENGINE_TILED), give identical trajectories, moment sums and per-sweep
    series, across split runs; and both equal the oracle.  The shapes cover
    whole tiles, tiles with a cut last row and column (every L that the side
    does not divide, in a launch of their own; odd L take two Philox calls for
    a site pair that straddles two, and cut rows are packed; cuts of 1 and 63
    sites: 65, 193, 257 and 127, 191, 255 on tiles of 64; below 32 the lattice is
    one cut tile of 32, its own neighbour: 3, 4, 5, 8, 16, 17, 31), slots merged in LDS
    (up to 6400: 320 = 10 x 10 tiles of 32) and in HBM (288 = 9 x 9 and 333 =
    11 x 11 tiles of 32, 333 = 6 x 6 and 352, 384 = 6 x 6 tiles of 64)."""
    from rbim._hip import ENGINE_TILED
    from rbim.engine import Batch

    specs = _specs(Ls, seed=12, temps=(1.3, 2.27, 3.1))
    n_therm, spb, n_bins = 5, 4, 3
    total = n_therm + spb * n_bins
    out = []
    for engine, per_launch, tile in [(PLAIN, 3, 0), (ENGINE_TILED, 4, 0), (PLAIN, 1, 0),
                                     (PLAIN, 2, 32), (PLAIN, 2, 64)]:
        with Batch(specs, n_therm=n_therm, n_bins=n_bins, sweeps_per_bin=spb, engine=engine,
                   sweeps_per_launch=per_launch, tile=tile, series=spb * n_bins) as batch:
            batch.run(2)   # the first run ends inside thermalisation
            batch.run(7)   # this one inside a bin
            batch.run()
            out.append((batch.moments(), batch.spins(), batch.series(), batch.info()))
    assert out[0][3].tiled_groups >= 1
    for acc, spins, series, _ in out[1:]:
        assert np.array_equal(acc, out[0][0])
        assert np.array_equal(series, out[0][2])
        for a, b in zip(spins, out[0][1]):
            assert np.array_equal(a, b)
    for i, spec in enumerate(specs[::3]):
        s_ref, acc_ref = _oracle_trajectory(coracle, spec, total, n_therm, spb, n_bins)
        k = 3 * i
        assert np.array_equal(out[0][1][k], s_ref), f"L={spec.L}"
        assert np.array_equal(out[0][0][k], acc_ref), f"L={spec.L}"
