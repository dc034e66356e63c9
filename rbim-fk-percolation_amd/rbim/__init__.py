"""FK cluster percolation in the 2D +-J random-bond Ising model, on MI355X.

Drop-in for the reference package ``rbim`` (``src/rbim/__init__.py:22-67``):
the same names, with every sweep, cluster labelling and measurement running
in hand-written gfx950 kernels (``librbim_hip.so``, C ABI in
``include/rbim_hip.h``) loaded through ctypes.  Disorder handling, the
estimators and file I/O stay on the host, as in the reference.
"""

from rbim.lattice import Lattice, build_lattice, couplings_from_disorder
from rbim.disorder import (
    generate_disorder,
    load_disorder,
    save_disorder,
    frustrated_plaquettes,
    frustration_density,
)
from rbim.clusters import (
    activate_bonds,
    label_clusters,
    cluster_statistics,
    wraps_along,
    spans_along,
)
from rbim.swendsen_wang import SweepResult, sw_sweep
from rbim.simulation import (
    RealizationResult,
    derive_observables,
    run_single_realization,
    run_temperature_scan,
)

__version__ = "0.1.0"

__all__ = [
    "Lattice",
    "build_lattice",
    "couplings_from_disorder",
    "generate_disorder",
    "load_disorder",
    "save_disorder",
    "frustrated_plaquettes",
    "frustration_density",
    "activate_bonds",
    "label_clusters",
    "cluster_statistics",
    "wraps_along",
    "spans_along",
    "SweepResult",
    "sw_sweep",
    "RealizationResult",
    "derive_observables",
    "run_single_realization",
    "run_temperature_scan",
]
