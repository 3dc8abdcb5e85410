"""MI355X backend for PolyFields.jl's SCFT hot path: a reference-shaped host API over
the C ABI of ``libpolyfields_hip.so`` (include/polyfields_hip.h)."""
from ._lib import LIB_PATH, PolyFieldsError, lib  # noqa: F401
from .api import (CANONICAL, GRAND, AndersonUpdater, Batch, Compressibility, DomainUpdater,  # noqa: F401
                  EdwardsInteraction, EulerUpdater, FieldSystem, FloryInteraction,
                  Homopolymer, Diblock, MomentumUpdater, Monomer, Multiblock, Point,
                  SCFTOptions, SCFTResults, UnitCell, chain_grid, fieldinit, scft, schedule_reset, schedule_set,
                  set_interaction, updater_with_system)
from .fileio import interpolate_grid, loadfields, savefields  # noqa: F401,E402
from .adjoint import (FUNCTIONALS, FhelmFunction, FhelmGrad, density_vjp,  # noqa: F401,E402
                      fhelm_grad, set_block_lengths)
