"""Reverse-mode derivative of the free energy on the GPU (``pf_fhelm_grad``,
``pf_density_vjp``, ``pf_species_set_blocks`` in include/polyfields_hip.h).

The reference has no derivative code; what is differentiated is its forward path,
``density!(sys)`` (src/system/system.jl:266-289) followed by ``free_energy``
(src/system/system.jl:402-433).  Two functionals are named:

* ``"reference"`` -- ``free_energy`` as the reference computes it (translational term
  phi (phi - 1) / N, the ``mu = phi`` line at system.jl:406-410);
* ``"textbook"`` -- the same with mu = log(phi / Q).

F is taken as a function of the fields as given (no mean subtraction), the cell
parameters, the Flory chi and Edwards u parameters, the species volume fractions and the
block lengths N_blk at a fixed contour grid (each block keeps its step count;
ds_b = v_b N_b / steps_b).  ``FhelmFunction`` wraps the same call as a
``torch.autograd.Function`` (the counterpart of a ChainRulesCore ``rrule``).
"""
import ctypes as C

import numpy as np

from ._lib import call, dptr
from .api import FloryInteraction, EdwardsInteraction, Point

FUNCTIONALS = {"reference": 0, "textbook": 1}


class FhelmGrad:
    """fhelm; dw {mid: grid}; dcell (nparams,); dchi {(a, b): value} per Flory pair;
    du {mid: value} per Edwards parameter; dphi (nspecies,); dN: per species, the list of
    dF/dN_blk (empty for a point species)."""

    def __repr__(self):
        return (f"FhelmGrad(fhelm={self.fhelm!r}, dcell={self.dcell!r}, dchi={self.dchi!r}, "
                f"dN={self.dN!r})")


def _chain_blocks(sys):
    return sum(sp.nblocks for sp in sys.species if not isinstance(sp, Point))


def _split_dN(sys, flat):
    out, k = [], 0
    for sp in sys.species:
        if isinstance(sp, Point):
            out.append([])
        else:
            out.append([float(x) for x in flat[k:k + sp.nblocks]])
            k += sp.nblocks
    return out


def fhelm_grad(sys, functional="reference"):
    """Free energy and its gradient at the current fields (runs ``density!``)."""
    if functional not in FUNCTIONALS:
        raise ValueError(f"functional must be one of {tuple(FUNCTIONALS)}")
    mids = sys.mids
    nm, nsp = len(mids), len(sys.species)
    npar = sys.cell.nparams
    dw = np.zeros((nm,) + sys.shape)
    dcell = np.zeros(npar)
    dchi = np.zeros((nm, nm))
    du = np.zeros(nm)
    dphi = np.zeros(nsp)
    ndN = _chain_blocks(sys)
    dN = np.zeros(max(ndN, 1))
    f = C.c_double()
    call("pf_fhelm_grad", sys._h, FUNCTIONALS[functional], C.byref(f), dptr(dw), dptr(dcell),
         npar, dptr(dchi), dptr(du), dptr(dphi), dptr(dN) if ndN else None, ndN)
    g = FhelmGrad()
    g.fhelm = f.value
    g.dw = {mid: dw[i] for i, mid in enumerate(mids)}
    g.dcell = dcell
    g.dchi, g.du = {}, {}
    for itx in sys.interactions:
        if isinstance(itx, FloryInteraction):
            for (a, b) in sorted(itx.chis):
                g.dchi[(a, b)] = float(dchi[mids.index(a), mids.index(b)])
        elif isinstance(itx, EdwardsInteraction):
            for m in sorted(itx.uvals):
                g.du[m] = float(du[mids.index(m)])
    g.dphi = dphi
    g.dN = _split_dN(sys, dN)
    return g


def density_vjp(sys, rho_bar, Q_bar):
    """Vector-Jacobian product of ``density!(sys)``: cotangents ``rho_bar`` {mid: grid}
    of the system densities and ``Q_bar`` (per species) of the partition functions ->
    (dw {mid: grid}, dcell, dN per species)."""
    mids = sys.mids
    rb = np.ascontiguousarray(np.stack([np.asarray(rho_bar[m], dtype=float).reshape(sys.shape)
                                        for m in mids]))
    qb = np.ascontiguousarray(np.asarray(Q_bar, dtype=float).reshape(len(sys.species)))
    dw = np.zeros((len(mids),) + sys.shape)
    dcell = np.zeros(sys.cell.nparams)
    ndN = _chain_blocks(sys)
    dN = np.zeros(max(ndN, 1))
    call("pf_density_vjp", sys._h, dptr(rb), dptr(qb), dptr(dw), dptr(dcell), len(dcell),
         dptr(dN) if ndN else None, ndN)
    return {m: dw[i] for i, m in enumerate(mids)}, dcell, _split_dN(sys, dN)


def set_block_lengths(sys, species, N_block):
    """Continuous block lengths for a chain species (object or index) on its fixed contour
    grid: each block keeps its step count and ds_b = v_b N_b / steps_b."""
    sp = sys.species[species] if isinstance(species, int) else species
    if isinstance(sp, Point):
        raise ValueError("block lengths are defined for chain species only")
    N = np.ascontiguousarray(np.asarray(N_block, dtype=float))
    call("pf_species_set_blocks", sys._h, sp._idx, dptr(N), len(N))
    vols = [sys.monomers[m].vol for m in sp.mid_block]
    sp.N_block = [float(x) for x in N]
    sp.N = float(sum(N))
    sp.Nref_block = [v * x for v, x in zip(vols, N)]
    sp.Nref = float(sum(sp.Nref_block))
    steps = [sp.begin_block[i + 1] - sp.begin_block[i] for i in range(sp.nblocks)]
    sp.ds_block = [nr / st for nr, st in zip(sp.Nref_block, steps)]
    sp.f_block = [nr / sp.Nref for nr in sp.Nref_block]


def _flory(sys):
    for itx in sys.interactions:
        if isinstance(itx, FloryInteraction):
            return itx
    return None


def _make_function():
    import torch

    class FhelmFunction(torch.autograd.Function):
        """``F = FhelmFunction.apply(sys, functional, w, cell, chi)``: w a float64 tensor
        (nmon, Nz, Ny, Nx) in ascending monomer id, cell (nparams,), chi (npairs,) in
        sorted Flory pair order.  The forward pass writes them into ``sys`` and runs the
        GPU gradient; backward scales the stored gradient."""

        @staticmethod
        def forward(ctx, sys, functional, w, cell, chi):
            wn = w.detach().cpu().numpy()
            for i, m in enumerate(sys.mids):
                sys.fields[m] = np.ascontiguousarray(wn[i])
            sys.cell.params = cell.detach().cpu().numpy().astype(float)
            itx = _flory(sys)
            pairs = sorted(itx.chis) if itx is not None else []
            for k, pr in enumerate(pairs):
                itx.set_interaction(pr[0], pr[1], float(chi[k]))
            g = fhelm_grad(sys, functional)
            dw = torch.from_numpy(np.stack([g.dw[m] for m in sys.mids]))
            dc = torch.from_numpy(np.array(g.dcell, dtype=float))
            dx = torch.tensor([g.dchi[pr] for pr in pairs], dtype=torch.float64)
            ctx.save_for_backward(dw, dc, dx)
            return torch.tensor(g.fhelm, dtype=torch.float64)

        @staticmethod
        def backward(ctx, go):
            dw, dc, dx = ctx.saved_tensors
            return None, None, go * dw, go * dc, go * dx

    return FhelmFunction


_FN = None


def FhelmFunction():
    """The ``torch.autograd.Function`` class (torch is imported on first use)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN
