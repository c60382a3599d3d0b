"""MI355X-native spectral-transform core of JGCM (CliMA/IdealizedSpectralGCM.jl).

Host-side mirror of the reference's Julia API over the C ABI of
``libjgcm_hip.so`` (include/jgcm.h).  Import through
``__graft_entry__.load_package()``, which registers this directory as the
module ``jgcm_amd`` (the directory name is not a Python identifier).
"""
from ._lib import JGCMError, lib
from .mesh import (Spectral_Spherical_Mesh, Trans_Spherical_To_Grid, Trans_Grid_To_Spherical,
                   Trans_Grid_To_Fourier, Divide_By_Cos, Multiply_By_Cos, Vor_Div_From_Grid_UV,
                   Compute_Alpha_Operator, UV_Grid_From_Vor_Div, Compute_Ucos_Vcos_From_Vor_Div,
                   Apply_Laplacian, Compute_Gradient_Cos, Compute_Gradients,
                   Add_Horizontal_Advection, Area_Weighted_Global_Mean,
                   Trans_Spherical_To_Grid_Batch, Trans_Grid_To_Spherical_Batch, BATCH_DIVCOS,
                   BATCH_VORDIV, INV_GRAD_LAMBDA, INV_GRAD_THETA, INV_UCOS, INV_VCOS)
from .time_integrator import (Filtered_Leapfrog, Update_Init_Step, Get_Δt, Get_ξ,
                              Compute_Spectral_Damping, Filtered_Leapfrog_step, Damped_Leapfrog,
                              Damped_Leapfrog_Batch)
from .dyn_data import Atmo_Data, Compute_Abs_Vor, Dyn_Data, Time_Advance
from .dynamics import (Barotropic_Dynamics, Shallow_Water_Physics, Implicit_Correction,
                       Shallow_Water_Dynamics, Shallow_Water_Step)
from .graph import Step_Graph
from .initial_conditions import Galewsky_Jet, Rossby_Haurwitz_Wave4
from .adjoint import Barotropic_Adjoint, Held_Suarez_Adjoint, Shallow_Water_Adjoint
from .output import Output_Manager, Update_Output, Finalize_Output, Zonal_Mean
from .primitive import (Vert_Coordinate, Pressure_Variables, Compute_Pressures_And_Heights,
                        Compute_Geopotential, Vert_Advection, Four_In_One, Semi_Implicit_Solver,
                        Implicit_Correction_HS, Spectral_Dynamics_Physics, Spectral_Dynamics,
                        Atmosphere_Update, Get_Topography, Spectral_Initialize_Fields)

__all__ = [
    "JGCMError", "lib", "Spectral_Spherical_Mesh", "Trans_Spherical_To_Grid",
    "Trans_Grid_To_Spherical", "Trans_Grid_To_Fourier", "Divide_By_Cos", "Multiply_By_Cos",
    "Vor_Div_From_Grid_UV", "Compute_Alpha_Operator", "UV_Grid_From_Vor_Div",
    "Compute_Ucos_Vcos_From_Vor_Div", "Apply_Laplacian", "Compute_Gradient_Cos",
    "Compute_Gradients", "Add_Horizontal_Advection", "Area_Weighted_Global_Mean",
    "Trans_Spherical_To_Grid_Batch", "Trans_Grid_To_Spherical_Batch", "BATCH_DIVCOS",
    "BATCH_VORDIV", "INV_GRAD_LAMBDA", "INV_GRAD_THETA", "INV_UCOS", "INV_VCOS",
    "Filtered_Leapfrog", "Update_Init_Step", "Get_Δt", "Get_ξ", "Compute_Spectral_Damping",
    "Filtered_Leapfrog_step", "Damped_Leapfrog", "Damped_Leapfrog_Batch", "Atmo_Data", "Compute_Abs_Vor", "Dyn_Data",
    "Time_Advance", "Barotropic_Dynamics", "Shallow_Water_Physics", "Implicit_Correction",
    "Shallow_Water_Step",
    "Shallow_Water_Dynamics", "Step_Graph", "Barotropic_Adjoint",
    "Shallow_Water_Adjoint", "Held_Suarez_Adjoint", "Galewsky_Jet", "Rossby_Haurwitz_Wave4",
]
