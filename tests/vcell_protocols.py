"""Variable-cell fixed-point protocols shared by the CPU floor test and the GPU parity
test (the lamellar diblock melt of the reference README: 64 x 1 x 1 grid, N = 100,
fA = 0.5, chi = 0.2, Ns = 100, Anderson nhist = 10, DomainUpdater nskip = 5, nper = 20).

SEEDED: fields 0.1 N(0,1) drawn as the reference draws them for seed 12345, cell L = 10.
    The reference's forward-difference cell Jacobian (dx = 1e-8, src/updaters/domain.jl:
    101) and its Re*Re stress (src/species/multiblock.jl:279-282), which depends on the
    lamellae's position in the cell, make the converged state depend on last-bit rounding
    along the way: the converged fhelm and L move by about 4e-9 and 3e-8 relative when
    the oracle's transforms are perturbed by one ulp (measured by the CPU floor test).
COSINE: fields +-cos(2 pi x) (centrosymmetric), cell L = 16 (near the optimum).  The
    lamellae stay where they start, and the converged fhelm and L are reproducible to
    about 1e-12 relative under the same perturbation.

Tolerances: ftol / stol as given; translation-invariant density comparisons use the
Fourier amplitudes |rho_hat(k)|.
"""
import numpy as np

from helpers import cos_fields, diblock_system

SEEDED = dict(L=10.0, seed=12345, nskip=5, nper=20, ftol=1e-12, stol=1e-14, nsteps=400)
COSINE = dict(L=16.0, seed=None, nskip=5, nper=20, ftol=1e-11, stol=1e-13, nsteps=400)


def run(mod, proto, **kw):
    """scft! of the protocol on ``mod`` (the oracle or the HIP package); returns the
    results, the cell parameter and the densities."""
    sys_, cell, _ = diblock_system(mod, (64, 1, 1), 1, "lamellar", (proto["L"],), chi=0.2, **kw)
    if proto["seed"] is None:
        mod.fieldinit(sys_, fields=cos_fields(sys_.shape))
    else:
        mod.fieldinit(sys_, seed=proto["seed"], scale=0.1)
    r = mod.scft(sys_, mod.AndersonUpdater(nhist=10),
                 domain=mod.DomainUpdater(nskip=proto["nskip"], nper=proto["nper"]),
                 opts=mod.SCFTOptions(nsteps=proto["nsteps"], ftol=proto["ftol"],
                                      stol=proto["stol"]))
    rho = {m: np.array(sys_.density[m], dtype=float) for m in (1, 2)}
    return r, float(cell.params[0]), rho


def compare(a, b):
    """Relative differences of two runs: fhelm, L, pointwise density (absolute) and
    density Fourier amplitudes (relative to the largest)."""
    (ra, La, da), (rb, Lb, db) = a, b
    dr = max(float(np.abs(da[m] - db[m]).max()) for m in (1, 2))
    amp = 0.0
    for m in (1, 2):
        fa, fb = np.abs(np.fft.rfft(da[m].ravel())), np.abs(np.fft.rfft(db[m].ravel()))
        amp = max(amp, float(np.abs(fa - fb).max() / fa.max()))
    return {"fhelm": abs(ra.fhelm - rb.fhelm) / abs(rb.fhelm), "L": abs(La - Lb) / abs(Lb),
            "rho": dr, "rho_amp": amp}
