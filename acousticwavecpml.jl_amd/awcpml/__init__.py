"""awcpml -- MI355X (gfx950) CPML acoustic time stepping for AcousticWaveCPML.jl.

Public surface (mirrors the reference's exports and solver entry points):
  Sources, Receivers, rickersource1D, calc_Kab_CPML   (src/utils.jl)
  solve2D, solve3D      (src/shared/acoustic_{2D,3D}_xPU.jl)
  Plan                  the C-ABI plan (multi-shot, FP64/FP32, graph capture)
  Batch                 several plans (uneven grids) advanced by one launch per step
  autotune              launch-shape autotuner with a JSON cache
  fwi                   FWI gradients through the GPU adjoint (misfit_gradient,
                        velocity_gradient, solve_traces: a torch.autograd.Function)
"""
from .host import Receivers, Sources, calc_Kab_CPML, rickersource1D, setup  # noqa: F401
from .solver import Batch, Plan, solve2D, solve3D  # noqa: F401
from .autotune import autotune, autotune_batch  # noqa: F401
from . import _lib  # noqa: F401
from . import fwi  # noqa: F401

__all__ = ["Sources", "Receivers", "rickersource1D", "calc_Kab_CPML", "setup", "Plan",
           "Batch", "autotune", "autotune_batch", "solve2D", "solve3D", "fwi"]
