"""litepgo_amd: litepgo's volumetric FEM hot path (neo-Hookean Tet P1 energy, gradient, Hessian and CSR assembly)
on AMD Instinct MI355X (gfx950), behind a C ABI (include/litepgo_fem.h)."""
from .batch import BatchedFemEnergy
from .energy import HessianKind
from .fem import VolumetricFemEnergy, kuhn_grid, lame_parameters
from .veg import VegMesh, read_veg, write_veg

__all__ = ["BatchedFemEnergy", "HessianKind", "VegMesh", "VolumetricFemEnergy", "kuhn_grid", "lame_parameters",
           "read_veg", "write_veg"]
