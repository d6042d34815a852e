"""tarang_amd -- MI355X (gfx950) backend for Tarang.jl's pseudo-spectral timestep.

The public names mirror Tarang.jl (src/public_api.jl:19-48); Julia's `!` suffix
is spelled with a trailing underscore.
"""
from ._lib import TarangError, LIB_PATH, load as load_library
from .core import (CPU, GPU, CartesianCoordinates, ChannelDomain, ChannelDomain3D, ChebyshevT, Coordinate, Distributor,
                   Domain, GPUTransformUnsupported, PeriodicDomain, RealFourier, ScalarField, VectorField,
                   backward_transform_, ensure_layout_, forward_transform_, get_grid_coordinates, padded_multiply, set_)
from .checkpoint import load_field_, load_state_, save_field, save_state
from .derivatives import evaluate_chebyshev_derivative_, gpu_chebyshev_deriv_
from .flow_tools import CFL, add_diffusivity_, add_velocity_, compute_timestep, grid_spacing
from .matsolvers import BatchedDenseLU, batched_factor_, batched_solve_, device_array
from .operators import Expr, derivative_basis, grad, lift, unit_vector_fields
from .problems import IVP, add_bc_, add_equation_, add_parameters_
from .solvers import (RK111, RK222, RK443, SBDF1, SBDF2, DiagonalIMEX_RK222, DiagonalIMEX_RK443, DiagonalIMEX_SBDF2,
                      InitialValueSolver, run_, step_)

__all__ = [
    "CPU", "GPU", "CartesianCoordinates", "ChannelDomain", "ChannelDomain3D", "ChebyshevT", "Coordinate",
    "Distributor", "Domain", "GPUTransformUnsupported", "PeriodicDomain", "RealFourier", "ScalarField", "VectorField",
    "backward_transform_", "ensure_layout_", "forward_transform_", "get_grid_coordinates", "padded_multiply", "set_",
    "IVP", "add_bc_", "add_equation_", "add_parameters_", "RK111", "RK222", "RK443", "SBDF1", "SBDF2",
    "DiagonalIMEX_RK222", "DiagonalIMEX_RK443", "DiagonalIMEX_SBDF2", "InitialValueSolver", "run_", "step_",
    "TarangError", "LIB_PATH", "load_library", "Expr", "derivative_basis", "grad", "lift", "unit_vector_fields",
    "evaluate_chebyshev_derivative_", "gpu_chebyshev_deriv_", "BatchedDenseLU", "batched_factor_", "batched_solve_",
    "device_array", "CFL", "add_diffusivity_", "add_velocity_", "compute_timestep", "grid_spacing",
    "load_field_", "load_state_", "save_field", "save_state",
]
