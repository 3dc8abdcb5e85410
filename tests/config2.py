"""The 64^3 cubic diblock workload: N = 100, fA = 0.35, bA = bB = 1, Ns = 100 (102 after
chain_grid), chi N = 20, cubic cell a = 4 Rg = 4 sqrt(100/6) b; fields 0.1 N(0,1) from
numpy's default_rng(seed) plus a BCC {110} cosine bias of amplitude 1.0 (+ on A, - on
B); variable cell.

SCFT protocol: MomentumUpdater(:MOM, lam = 0.1) with DomainUpdater(nskip = 100,
nper = 10), ftol = stol = 1e-5 (src/updaters/momentum.jl, src/updaters/domain.jl,
src/simulate.jl).  From this seed the reference's Anderson mixing diverges (its step
size reaches 1 once the history fills, anderson.jl:103-110), with or without a first
phase of Euler or momentum steps, so the momentum updater is the protocol.
"""
import math
import warnings

import numpy as np

N = 100
FA = 0.35
CHI = 20.0 / N
A_CELL = 4.0 * math.sqrt(N / 6.0)
SEED = 2024
PROTO = dict(lam=0.1, nskip=100, nper=10, ftol=1e-5, stol=1e-5)


def seeded_fields(shape, seed=SEED):
    """0.1 N(0,1) + the BCC {110} bias cos 2pi(x+-y) + cos 2pi(y+-z) + cos 2pi(x+-z), / 6."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx,
                          indexing="ij")
    tp = 2 * np.pi
    bias = (np.cos(tp * (x + y)) + np.cos(tp * (x - y)) + np.cos(tp * (y + z)) +
            np.cos(tp * (y - z)) + np.cos(tp * (x + z)) + np.cos(tp * (x - z))) / 6.0
    return {1: 0.1 * rng.standard_normal(shape) + bias,
            2: 0.1 * rng.standard_normal(shape) - bias}


def build(mod, grid=64, a=A_CELL, **kw):
    cell = mod.UnitCell(3, "cubic", a)
    A, B = mod.Monomer(1), mod.Monomer(2)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        ch = mod.Diblock(A, B, N, FA, 1.0, 1.0, 100)
    itx = mod.FloryInteraction()
    itx.set_interaction(1, 2, CHI)
    s = mod.FieldSystem((grid, grid, grid), cell, monomers=[A, B], **kw)
    s.add_species(ch, 1.0)
    s.add_interaction(itx)
    mod.fieldinit(s, fields=seeded_fields(s.shape))
    return s, cell, ch


def run(mod, s, nsteps, nskip=PROTO["nskip"], nper=PROTO["nper"], ftol=PROTO["ftol"],
        stol=PROTO["stol"]):
    return mod.scft(s, mod.MomentumUpdater(method="MOM", lam=PROTO["lam"]),
                    domain=mod.DomainUpdater(nskip=nskip, nper=nper),
                    opts=mod.SCFTOptions(nsteps=nsteps, ftol=ftol, stol=stol))


def amplitudes(rho):
    """Translation-invariant comparison: |rho_hat(k)| over the half spectrum."""
    return np.abs(np.fft.rfftn(np.asarray(rho, dtype=float)))
