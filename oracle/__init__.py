"""FP64 CPU restatement of PolyFields.jl's SCFT hot path -- TEST INFRASTRUCTURE ONLY.

This package is the parity oracle.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker or as the
reported CPU baseline.  The product path (``polyfields.jl_amd/``) never imports it and
has no CPU fallback.

Every function cites the reference file:line it restates (paths relative to the
PolyFields.jl source tree).  Arrays use the reference's memory order: a Julia
``Array{Float64,3}`` of size (Nx, Ny, Nz), column-major with x fastest, is held here as a
C-ordered numpy array of shape (Nz, Ny, Nx).  The half spectrum is taken along x, the
reference's first dimension (``src/utils.jl:9``).

Parity pins (what checks each function against something other than itself) are listed
in ``tests/test_oracle_*.py``.  ``fieldinit`` restates Julia 1.5's seeded draws
(``julia_rng``), which the reference's own seeded known answers pin
(test/test_system.jl:29-34, test/test_updaters.jl:24,30; tests/test_oracle_julia_rng.py).
"""

from .cell import UnitCell, make_basis, basis_matrix, basis_dimensions  # noqa: F401
from .chain import (Monomer, Multiblock, Diblock, Homopolymer, Point,  # noqa: F401
                    chain_grid)
from .system import (FieldSystem, FloryInteraction, EdwardsInteraction,  # noqa: F401
                     Compressibility, fieldinit, set_fft_workers)
from .updaters import (AndersonUpdater, DomainUpdater, EulerUpdater,  # noqa: F401
                       MomentumUpdater)
from .simulate import SCFTOptions, SCFTResults, scft, isconverged  # noqa: F401
