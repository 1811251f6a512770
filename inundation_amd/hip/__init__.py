"""The ``hip_resident`` backend: ctypes binding and solver class."""

from .solver import SWESolverHipResident, available, device_mesh_arrays

__all__ = ["SWESolverHipResident", "available", "device_mesh_arrays"]
