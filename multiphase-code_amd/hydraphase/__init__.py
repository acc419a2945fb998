"""hydraphase: MI355X (gfx950) implementation of the HydraPhase two-phase
tetrahedral finite-volume step, behind the C ABI of include/hydra_hip.h."""
from .solver import Solver  # noqa: F401
