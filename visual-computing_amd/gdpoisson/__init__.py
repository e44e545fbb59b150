"""gdpoisson: MI355X-native multigrid Poisson integration for gradient-domain imaging.

Drop-in GPU replacement for the Poisson path of the gradient-domain package
(Giankaqui/visual-computing, ``gradient-domain/src/gradient_domain``): the same
public names (``solve_system``, ``SOLVERS``, ``SolverReport``, ``MultigridOptions``,
``MultigridReport``, ``v_cycle``, ``solve``, ``solve_neumann``, ``gradient``,
``divergence``, ``laplacian``, ``fold_boundary``) backed by hand-written HIP kernels
in ``libgdpoisson.so`` (C ABI: ``include/gdpoisson.h``).  Batched plans
(:class:`Plan`) solve many grids per launch.  The compositing (``seamless_clone``,
``GuidanceField``, ``solve_dirichlet``, ``solve_masked``, ``texture_flatten``,
``illumination_change``) and tone-mapping (``tone_map``,
``attenuation_field``, ``ToneMapConfig``) callers of the solve run on the device too.
"""

from .multigrid import MultigridOptions, MultigridReport, solve, v_cycle
from .operators import (divergence, fold_boundary, gradient, laplacian, neumann_laplacian,
                        sparse_laplacian)
from .autotune import autotune
from .plan import Plan, PlanOptions
from .poisson import (GuidanceField, illumination_change, seamless_clone, solve_dirichlet,
                      solve_masked, texture_flatten)
from . import hdr
from .hdr import ToneMapConfig, attenuation_field, tone_map
from .solvers import SOLVERS, SolverReport, solve_neumann, solve_system
from .benchmark import BenchmarkRecord, BenchmarkTable, manufactured_problem, solver_scaling

__version__ = "0.1.0"

__all__ = [
    "MultigridOptions",
    "MultigridReport",
    "Plan",
    "GuidanceField",
    "seamless_clone",
    "texture_flatten",
    "illumination_change",
    "solve_dirichlet",
    "solve_masked",
    "ToneMapConfig",
    "attenuation_field",
    "hdr",
    "tone_map",
    "autotune",
    "PlanOptions",
    "SOLVERS",
    "SolverReport",
    "divergence",
    "fold_boundary",
    "gradient",
    "laplacian",
    "neumann_laplacian",
    "solve",
    "solve_neumann",
    "solve_system",
    "v_cycle",
    "sparse_laplacian",
    "manufactured_problem",
    "solver_scaling",
    "BenchmarkRecord",
    "BenchmarkTable",
]
