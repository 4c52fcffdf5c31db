"""3D solver: ``SolverConfig3D`` / ``FlowSolver3D`` API on the MI355X step."""
from .config import AnalyticBody, DyeRake, ImmersedMesh, SolverConfig3D

__all__ = ["AnalyticBody", "DyeRake", "ImmersedMesh", "SolverConfig3D"]
