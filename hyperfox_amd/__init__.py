"""hyperfox_amd: MI355X-native HDG element-by-element hot path.

The compute path is libhfx_hip.so (HIP kernels for gfx950 behind a C ABI,
declared in include/hfx_hip.h).  This package is a thin ctypes layer over it
for the tests and the benchmark; it has no CPU fallback: importing
``hyperfox_amd.lib`` raises if the shared library has not been built.
"""
from ._lib import lib, LIB_PATH, HfxError  # noqa: F401
from .hdg import Mesh, HDGProblem, Batch, refel_sizes, refel_tables, cubature, ilu_analyse  # noqa: F401
from .hdg import RK_NAMES, RK_CUSTOM, rk_table  # noqa: F401
from .linalg import HipLinAlgebra, HipOpts  # noqa: F401
from .burgers import BurgersProblem, MultiProblem, NGammaProblem  # noqa: F401

MODEL_LAPLACE = 0
MODEL_DIFFUSION_SOURCE = 1
MODEL_CDR_SOURCE = 2
PC_NONE, PC_JACOBI, PC_BJACOBI, PC_LEARNED, PC_ILU0, PC_ILU0_NATURAL, PC_ILU0_F32 = 0, 1, 2, 3, 4, 5, 6
ILU_MULTICOLOUR, ILU_NATURAL = 0, 1
KRYLOV_FGMRES, KRYLOV_BICGSTAB = 0, 1
TS_NONE, TS_BEULER, TS_FEULER, TS_RK = 0, 1, 2, 3
BC_DIRICHLET, BC_INTEGRATED_DIRICHLET = 0, 1
HDG_IMPLICIT, HDG_WEXPLICIT, HDG_SEXPLICIT = 0, 1, 2
