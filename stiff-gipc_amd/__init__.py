"""MI355X-native Newton linear-solve path of StiffGIPC: 3x3 upper-BCOO Hessian
assembly, SRBK SpMV, MAS preconditioner and PCG on gfx950, behind the C ABI
declared in ``include/sgipc.h``."""
from . import learned, mesh_io, partition, scenes  # noqa: F401
from ._lib import LIB_PATH, SgipcError, exported_symbols  # noqa: F401
from .linear_system import (FP32_IR, FP32_MAS, FP64, STOP_REF_RZ, STOP_TRUE_RESIDUAL, Context,  # noqa: F401
                            ContactDetector, Converter, DevArray, GlobalLinearSystem, LinearSystem, MASPreconditioner, PCGSolver,
                            SolveResult, SpMat, Spmv, pcg_opts, upload_scene)
