"""The C ABI library without a GPU: it loads, exports every function declared
in include/rbim_hip.h with the signature the ctypes binding uses, validates
arguments on the host, and refuses to run (loudly) when no GPU is present."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "rbim_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"^\s*(?:int|const char\*)\s+(rbim_\w+)\s*\(", text, flags=re.M))


def test_library_exports_every_declared_symbol():
    from rbim import _hip

    lib = _hip.lib()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert names == set(_hip.SIGNATURES)


def test_abi_version():
    from rbim import _hip

    assert _hip.lib().rbim_abi_version() == 3


def test_descriptor_validation_on_host():
    from rbim import _hip
    from rbim.engine import LatticeSpec, footprint

    spec = LatticeSpec(L=32, J_h=np.ones((32, 32), np.int8), J_v=np.ones((32, 32), np.int8),
                       beta=0.5, seed=1, lattice_id=0)
    lds = footprint([spec] * 10, n_bins=10)
    tiled = footprint([spec] * 10, n_bins=10, engine=_hip.ENGINE_PLAIN)
    assert tiled > lds > 10 * 32 * 32 * 2
    big = LatticeSpec(L=1024, J_h=np.ones((1, 1), np.int8), J_v=np.ones((1, 1), np.int8),
                      beta=0.5, seed=1, lattice_id=0)
    # edge-slot layout (spins, couplings, two planes of 16-bit state words, couplings
    # packed tile by tile, slot forest in HBM): ~8 B per site;
    # multi-kernel tiled layout: ~27 B per site
    assert 1024 * 1024 * 7 < footprint([big], n_bins=1) < 1024 * 1024 * 10
    assert footprint([big], n_bins=1, engine=_hip.ENGINE_TILED) > 1024 * 1024 * 25
    with pytest.raises(ValueError, match="at least 3"):
        footprint([LatticeSpec(L=2, J_h=None, J_v=None, beta=1, seed=0, lattice_id=0)], n_bins=1)
    with pytest.raises(ValueError, match="LDS engine"):
        footprint([big], n_bins=1, engine=_hip.ENGINE_BATCHED)
    with pytest.raises(ValueError, match="n_bins"):
        footprint([spec], n_bins=0)


def test_fails_loudly_without_a_gpu():
    from rbim import _hip

    try:
        _hip.context()
    except _hip.EngineUnavailable as exc:
        assert "device" in str(exc)
    else:
        pytest.skip("a GPU is present here")
    import rbim

    lattice = rbim.build_lattice(4)
    with pytest.raises(_hip.EngineUnavailable):
        rbim.label_clusters(np.zeros(lattice.n_bonds, bool), lattice)


def test_footprint_by_layout():
    """rbim_batch_footprint is host arithmetic (no GPU): the edge-slot layout
    holds less than the multi-kernel tiled layout, the LDS layout least, and a
    series ring adds n_lattices x capacity x sizeof(rbim_sweep_stats)."""
    import numpy as np

    from rbim import _hip
    from rbim.engine import LatticeSpec, footprint

    def specs(L, n):
        J = np.ones((L, L), np.int8)
        return [LatticeSpec(L=L, J_h=J, J_v=J, beta=0.4, seed=1, lattice_id=i) for i in range(n)]

    big = specs(256, 8)
    fused = footprint(big, n_bins=10, engine=_hip.ENGINE_PLAIN)
    tiled = footprint(big, n_bins=10, engine=_hip.ENGINE_TILED)
    assert 0 < fused < tiled
    small = specs(32, 8)
    assert footprint(small, n_bins=10) < footprint(small, n_bins=10, engine=_hip.ENGINE_PLAIN)
    ring = footprint(small, n_bins=10, series=100) - footprint(small, n_bins=10)
    assert ring == 8 * 100 * ctypes.sizeof(_hip.SweepStats)
    with pytest.raises(ValueError):
        footprint(small, n_bins=10, series=-1)
