"""MI355X-native executor for TensorOrder's contraction phase.

A new ``tensor_library`` choice behind the reference's plugin interface
(``src/tensor_network/tensor_apis/base_api.py``), plus the host-side mirror of the
structures that interface reads: tensor networks, contraction trees, sliced plans and
slicers.  The HIP library does all contraction arithmetic; PyTorch only allocates
device memory.
"""

from .api import BaseTensorAPI, OutOfMemoryError
from .network import TensorNetwork, Tensor, OrTensor, VariableTensor, cnf_count, ALL_CONSTRUCTIONS
from .plan import SlicedExecutionPlan, ALL_SLICERS
from .formula import Formula, WeightFormat


def _hip():
    from .hip_api import HipAPI

    return HipAPI()


def _hip_f32():
    from .hip_api import HipF32API

    return HipF32API()


# Zero-argument factories, as the reference's registry holds (tensor_network/__init__.py:12-16).
ALL_APIS = {"hip": _hip, "hip-f32": _hip_f32}

__all__ = [
    "BaseTensorAPI", "OutOfMemoryError", "TensorNetwork", "Tensor", "OrTensor", "VariableTensor",
    "cnf_count", "ALL_CONSTRUCTIONS", "SlicedExecutionPlan", "ALL_SLICERS", "Formula",
    "WeightFormat", "ALL_APIS",
]
