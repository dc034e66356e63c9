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

    assert _hip.lib().rbim_abi_version() == 2


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
    assert footprint([big], n_bins=1) > 1024 * 1024 * 25  # tiled state, ~27 B per site
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
