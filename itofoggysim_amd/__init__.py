"""MI355X-native multi-sphere scattering matvec (the ItoFoggySim hot path).

The compute path is libifs.so (HIP, gfx950) behind the C ABI in include/ifs.h;
this package is the host-side mirror of the reference's solver interface.
"""
from ._lib import IfsError, lib  # noqa: F401
from .cluster import make_cluster, config_cluster  # noqa: F401
from .solver import ClusterBatch, MultiSphereSolver  # noqa: F401
from .serve import Server, plan_request, poisson_workload, run_workload, water_index  # noqa: F401

__version__ = "0.1.0"
