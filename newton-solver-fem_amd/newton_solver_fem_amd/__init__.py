"""MI355X-native SolverFEM substep: hand-written HIP kernels behind a C-ABI.

Public names mirror the reference package (``newton_solver_fem``):
``SolverFEM`` (alias ``SolverLinearFEM``), ``FEMBuilder`` and
``LinearFEMData``; the Newton-side containers it needs (``ModelBuilder``,
``Model``, ``State``, ``Control``) are provided here.
"""

from .model import (
    PARTICLE_FLAG_ACTIVE,
    Contacts,
    Control,
    FEMBuilder,
    HashGrid,
    LinearFEMData,
    Model,
    ModelBuilder,
    State,
    build_hex_model,
    build_tet_model,
    hex_reference_data,
    merge_tet_scenes,
)

__all__ = [
    "PARTICLE_FLAG_ACTIVE",
    "Contacts",
    "Control",
    "FEMBuilder",
    "HashGrid",
    "LinearFEMData",
    "Model",
    "ModelBuilder",
    "PreparedFEMStep",
    "State",
    "SolverFEM",
    "SolverLinearFEM",
    "build_hex_model",
    "build_tet_model",
    "hex_reference_data",
    "merge_tet_scenes",
]


def __getattr__(name):
    if name in ("SolverFEM", "SolverLinearFEM"):
        from .solver import SolverFEM

        return SolverFEM
    if name == "PreparedFEMStep":
        from .prepared_step import PreparedFEMStep

        return PreparedFEMStep
    raise AttributeError(name)
