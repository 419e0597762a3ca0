"""MI355X-native forward solve (SIMPLE + SST k-omega) with the reference's
ForwardBatch interface. See the module docstrings for the reference
interfaces each part mirrors."""
from .config import (CONVERGED, DIVERGED, INVALID_PARAMETERS, STATUS_NAMES, UNCONVERGED, UNKNOWN,
                     FlowBoundaryConditions, InferenceParameterSet, ObservationOperator,
                     PatchBC, PreparedObservation, SolverSettings, Vec3)
from .mesh import Mesh, Patch, hill_bottom, make_hill_mesh

__all__ = ["Mesh", "Patch", "hill_bottom", "make_hill_mesh", "FlowBoundaryConditions",
           "InferenceParameterSet", "ObservationOperator", "PatchBC", "PreparedObservation",
           "SolverSettings", "Vec3", "STATUS_NAMES", "CONVERGED", "UNCONVERGED", "DIVERGED",
           "INVALID_PARAMETERS", "UNKNOWN"]
