"""MI355X (gfx950) solute-solvent energy pipeline of the energy-representation
method: the per-frame path of ERmod's engproc / realcal / reciprocal modules,
as hand-written HIP kernels behind the C-ABI of include/ermgpu.h."""
from .system import ErmSystem, Hist  # noqa: F401

__all__ = ["ErmSystem", "Hist"]
