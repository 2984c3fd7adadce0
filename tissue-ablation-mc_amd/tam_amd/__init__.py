"""Host-side mirror of the Tissue-Ablation-MC hot path over libtam.so (HIP, gfx950)."""
from . import lib, configs  # noqa: F401
from .lib import TissueGrid, Params, Pulse, MCStats, load  # noqa: F401
