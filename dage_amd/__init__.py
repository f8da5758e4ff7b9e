"""dage_amd -- MI355X (gfx950) implementation of DAGE's operator-application
hot path: the tensor-quadrature Coulomb3D/Helmholtz3D apply on the cube, the
grid-based FMM Coulomb potential and the bubbles inject.

The product is the C-ABI library dage_amd/libdage_ops.so (include/dage_ops.h);
this package is its host-side mirror of the reference's Operator3D API.
"""
from ._lib import (DAGE_OP_NO_CODA, DAGE_PREC_BF16X3, DAGE_PREC_BF16X6, DAGE_PREC_FP64,
                   DageError, build, load)
from .operators import (Bubbles, Context, Coulomb3D, CubeView, DeviceCube, GaussQuad, GBFMMCoulomb3D,
                        Grid3D, Helmholtz3D, HelmholtzBatch, Operator3D, interaction_matrix,
                        radial_grid_steps,
                        translation_matrix)

__all__ = [
    "Bubbles", "Context", "Coulomb3D", "CubeView", "DeviceCube", "GaussQuad", "GBFMMCoulomb3D", "Grid3D",
    "Helmholtz3D", "HelmholtzBatch", "Operator3D", "interaction_matrix", "radial_grid_steps", "translation_matrix",
    "DageError", "DAGE_OP_NO_CODA", "DAGE_PREC_FP64", "DAGE_PREC_BF16X6", "DAGE_PREC_BF16X3",
    "build", "load",
]
