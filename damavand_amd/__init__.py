"""MI355X-native statevector path of damavand (apply_method="gpu").

``Circuit`` mirrors the reference's PyO3 class; ``StateVector`` is a
device-resident state; ``AdjointExpectation`` (torch) and
``batch_forward_adjoint`` are the gradient and batching additions.  All of it
runs in libdamavand-gpu.so (hand-written HIP kernels for gfx950); build it with
``python -m damavand_amd.build``.
"""
from ._native import DamavandError, NoDeviceError, OutOfDeviceMemory  # noqa: F401
from .circuit import APPLY_METHODS, Circuit, StateVector, initialize_mpi  # noqa: F401
from .batch import (PackedBatch, batch_autotune, batch_forward_adjoint, batch_forward_adjoint_packed,  # noqa: F401
                    pack_circuits)

__all__ = ["Circuit", "StateVector", "initialize_mpi", "batch_forward_adjoint", "batch_forward_adjoint_packed",
           "batch_autotune", "pack_circuits", "PackedBatch", "DamavandError",
           "NoDeviceError", "OutOfDeviceMemory", "APPLY_METHODS"]
