"""2D solver: ``SolverConfig`` / ``FlowSolver`` API on the MI355X step."""
from .config import EmitterSpec, HeatSourceSpec, ImmersedBody, SolverConfig, ThermalConfig, WallBC

__all__ = ["EmitterSpec", "HeatSourceSpec", "ImmersedBody", "SolverConfig", "ThermalConfig", "WallBC"]
