"""cuteflow_amd: MI355X-native FP64 finite-volume St-Venant time step on
unstructured triangles, a drop-in for CuteFlow's time-step path on one
subdomain.  Compute runs in libcuteflow_hip.so (HIP, gfx950)."""
from .mesh import Mesh, read_text_mesh, write_text_mesh, read_boundary_table, write_boundary_table
from .params import Params, read_donnees, parse_namelist
from .solver import BatchSolver, Solver, autotune, initial_state, solve_finite_volumes

__all__ = [
    "Mesh", "read_text_mesh", "write_text_mesh", "read_boundary_table", "write_boundary_table",
    "Params", "read_donnees", "parse_namelist", "Solver", "BatchSolver", "autotune", "initial_state", "solve_finite_volumes",
]
