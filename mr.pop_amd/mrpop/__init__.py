"""MR.POP on MI355X: Python view of the C ABI (include/mrpop_c.h)."""
from .scenes import Problem, Scene  # noqa: F401
