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
                   Add_Horizontal_Advection, Area_Weighted_Global_Mean)

__all__ = [
    "JGCMError", "lib", "Spectral_Spherical_Mesh", "Trans_Spherical_To_Grid",
    "Trans_Grid_To_Spherical", "Trans_Grid_To_Fourier", "Divide_By_Cos", "Multiply_By_Cos",
    "Vor_Div_From_Grid_UV", "Compute_Alpha_Operator", "UV_Grid_From_Vor_Div",
    "Compute_Ucos_Vcos_From_Vor_Div", "Apply_Laplacian", "Compute_Gradient_Cos",
    "Compute_Gradients", "Add_Horizontal_Advection", "Area_Weighted_Global_Mean",
]
