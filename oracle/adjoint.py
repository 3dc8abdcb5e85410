"""Reverse-mode derivative of the free energy -- TEST INFRASTRUCTURE ONLY (FP64 restatement).

The reference has no derivative code; what is differentiated is its own forward path:
``density!`` of each species (src/species/multiblock.jl:125-231, src/species/point.jl:36-53)
with the RK2 Strang step (src/system/mde.jl:65-81), the pooling in ``density!(sys)``
(src/system/system.jl:266-289) and ``free_energy`` (src/system/system.jl:402-433).  Two
functionals are named explicitly:

* ``"reference"``: ``free_energy`` exactly as the reference computes it, including its
  translational term phi (phi - 1) / N (the ``mu = phi`` line, system.jl:406-410);
* ``"textbook"``: the same with mu = log(phi / Q), the translational entropy the
  reference's ``muphi!`` computes (system.jl:248-259) but ``free_energy`` does not use.

F is a function of the fields w (as given; nothing is mean-subtracted), the cell
parameters, the interaction parameters chi, the block lengths N_blk (at a fixed contour
grid: every block keeps its step count and ds_blk = v N_blk / (2 Ns_half)) and the
species volume fractions phi.  Canonical ensemble only.

The derivative is the adjoint of the discrete sweep.  The Strang operator
P = LW F^-1 LD F LW is real-symmetric, so the adjoint of q_{s+1} = P q_s runs the same
operator in the opposite direction, with the Simpson sources
(ds_b / (3 Q Nref)) w_s phi rbar q+_s injected at every slice.  Per step, with
a-bar the adjoint at the output slice, A = rfft(LW a-bar) and B = rfft(LW q_in):

    dF/dw(r)      += -(ds v / 2) (a-bar q_out + q_in P a-bar)
    dF/dtheta     += -(ds b^2 / 6n) sum_k m_k Re(conj(A) B) LD dksq_theta
    dF/dds (LD)   += -(b^2 / 6n)    sum_k m_k Re(conj(A) B) LD ksq

with m_k = 1 on the kx = 0 and kx = Nx/2 planes of the half spectrum and 2 elsewhere.
Pinned in tests/test_oracle_adjoint.py by central finite differences of the oracle's
own free energy in every argument, on 1-D, 2-D (hexagonal) and 3-D cells.
"""
import math

import numpy as np

from .chain import Point
from .system import CANONICAL, rfft3, irfft3


FUNCTIONALS = ("reference", "textbook")


def _check(sys, functional):
    if functional not in FUNCTIONALS:
        raise ValueError(f"functional must be one of {FUNCTIONALS}")
    if sys.ensemble != CANONICAL:
        raise ValueError("the free-energy derivative is defined for the canonical ensemble")
    if sys.mde != "RK2":
        raise ValueError("the free-energy derivative is defined for the RK2 propagator")


def _ftrans(phi, Q, Nref, functional):
    if phi <= 1e-8:
        return 0.0
    mu = phi if functional == "reference" else math.log(phi / Q)
    return phi * (mu - 1.0) / Nref


def free_energy(sys, functional="reference"):
    """F at the current fields: ``density!`` then ``free_energy`` (system.jl:402-433), the
    translational term per ``functional``.  Fields are used as given."""
    _check(sys, functional)
    sys.density_update()
    f = 0.0
    for isp, sp in enumerate(sys.species):
        f += _ftrans(sys.phi_species[isp], sp.Q, sp.Nref, functional)
    ff = 0.0
    for mid in sys.mids:
        ff -= np.sum(sys.density[mid] * sys.fields[mid])
    f += ff / sys.ngrid
    f += sum(itx.energy(sys) for itx in sys.interactions)
    return f


def set_block_lengths(sp, N_block):
    """Continuous block lengths at a fixed contour grid: Nref_b = v_b N_b,
    ds_b = Nref_b / (2 Ns_half_b) with Ns_half_b and the begin indices unchanged
    (src/species/base.jl:27-56 evaluated with the step counts held)."""
    steps = [(sp.begin_block[i + 1] - sp.begin_block[i]) for i in range(sp.nblocks)]
    sp.N_block = [float(x) for x in N_block]
    sp.Nref_block = [sp.vol_block[i] * sp.N_block[i] for i in range(sp.nblocks)]
    sp.Nref = sum(sp.Nref_block)
    sp.N = sum(sp.N_block)
    sp.ds_block = [sp.Nref_block[i] / steps[i] for i in range(sp.nblocks)]
    sp.f_block = [Nr / sp.Nref for Nr in sp.Nref_block]


def _mult(sys):
    nx = sys.dims[0]
    m = np.full(sys.kshape, 2.0)
    m[..., 0] = 1.0
    if nx % 2 == 0:
        m[..., nx // 2] = 1.0
    return m


def _simpson_weight(s, bgn, end):
    """Simpson weight of 1-based slice s inside block [bgn, end] (multiblock.jl:182-216)."""
    if s == bgn or s == end:
        return 1.0
    return 4.0 if (s - bgn) % 2 == 1 else 2.0


class ChainVJP:
    """Per-species cotangent outputs."""

    def __init__(self, nblk, nparams, shape):
        self.wbar = [np.zeros(shape) for _ in range(nblk)]   # dF/dw per block
        self.dtheta = np.zeros(nparams)
        self.dds_ld = np.zeros(nblk)                           # LD part of dF/dds_b


def _chain_vjp(sys, sp, phi, rbar, Qtot):
    """Adjoint sweeps of one chain species.  ``rbar``: dF/d(system density) per monomer;
    ``Qtot``: total dF/dQ (the 1/Q of the normalisation included)."""
    n = sys.ngrid
    Ns, Q, Nref = sp.Ns, sp.Q, sp.Nref
    cell = sys.cell
    mult = _mult(sys)
    out = ChainVJP(sp.nblocks, cell.nparams, sys.shape)
    nb = sp.nblocks
    homo = nb == 1

    def coef(s):
        """Source field C_s(r) = sum over blocks holding slice s of
        (ds_b / (3 Q Nref)) w^b_s phi rbar_{mon(b)}."""
        c = np.zeros(sys.shape)
        for b in range(nb):
            bgn, end = sp.begin_block[b], sp.begin_block[b + 1]
            if bgn <= s <= end:
                c += (sp.ds_block[b] / (3.0 * Q * Nref)) * _simpson_weight(s, bgn, end) * \
                    phi * rbar[sp.mid_block[b]]
        return c

    def ops(b):
        mon = sys.monomers[sp.mid_block[b]]
        ds, bb = sp.ds_block[b], sp.b_block[b]
        LW = np.exp(-(ds / 2.0) * mon.vol * sys.fields[mon.id])
        LD = np.exp(-(ds * bb ** 2 / 6.0) * cell.ksq)
        return LW, LD, ds, bb, mon.vol

    def step(b, abar, q_in, q_out, LWLD):
        LW, LD, ds, bb, vol = LWLD
        A = rfft3(LW * abar)
        B = rfft3(LW * q_in)
        g = mult * np.real(np.conj(A) * B) * LD
        for k in range(cell.nparams):
            out.dtheta[k] += -(ds * bb ** 2 / (6.0 * n)) * np.sum(g * cell.dksq[k])
        out.dds_ld[b] += -(bb ** 2 / (6.0 * n)) * np.sum(g * cell.ksq)
        Pa = LW * irfft3(A * LD, sys.shape)
        out.wbar[b] += -(ds * vol / 2.0) * (abar * q_out + q_in * Pa)
        return Pa

    q, qc = sp.q, sp.qc
    other = (lambda s: q[Ns + 1 - s]) if homo else (lambda s: qc[s - 1])
    fac = 2.0 if homo else 1.0
    tables = [ops(b) for b in range(nb)]

    # adjoint of q: from s = Ns+1 down to 1
    abar = Qtot / n + fac * coef(Ns + 1) * other(Ns + 1)
    for b in range(nb - 1, -1, -1):
        for s_in in range(sp.begin_block[b + 1] - 1, sp.begin_block[b] - 1, -1):
            Pa = step(b, abar, q[s_in - 1], q[s_in], tables[b])
            abar = Pa + fac * coef(s_in) * other(s_in)
    if homo:
        return out
    # adjoint of q+: from s = 1 up to Ns+1
    abar = coef(1) * q[0]
    for b in range(nb):
        for s_out in range(sp.begin_block[b], sp.begin_block[b + 1]):
            Pa = step(b, abar, qc[s_out], qc[s_out - 1], tables[b])
            abar = Pa + coef(s_out + 1) * q[s_out]
    return out


class FhelmGrad:
    """dF with respect to every argument.

    fhelm: F; dw: {mid: grid}; dcell: (nparams,); dchi: {(a, b): value} for each Flory
    pair; du: {mid: value} for each Edwards parameter; dN: per species, the list of
    dF/dN_blk (empty for a point species); dphi: (nspecies,)."""


def fhelm_grad(sys, functional="reference"):
    """Reverse-mode derivative of ``free_energy`` (see module docstring)."""
    _check(sys, functional)
    F = free_energy(sys, functional)              # runs density_update
    n = sys.ngrid
    sys.potentials_update()
    rbar = {}
    for mid in sys.mids:
        v = sys.monomers[mid].vol
        rbar[mid] = (sys.potentials[mid] / v - sys.fields[mid]) / n
    g = FhelmGrad()
    g.fhelm = F
    g.dw = {mid: -sys.density[mid] / n for mid in sys.mids}
    g.dcell = np.zeros(sys.cell.nparams)
    g.dN = []
    g.dphi = np.zeros(len(sys.species))
    for isp, sp in enumerate(sys.species):
        phi = sys.phi_species[isp]
        Q = sp.Q
        active = phi > 1e-8
        mu = (phi if functional == "reference" else math.log(phi / Q)) if active else 0.0
        Qbar = 0.0
        if functional == "textbook" and active:
            Qbar = -phi / (sp.Nref * Q)
        # <rbar, rho_p> summed over the species' monomers
        rr = sum(float(np.sum(rbar[mid] * sp.density[mid])) for mid in sp.mids)
        g.dphi[isp] = rr
        if active:
            g.dphi[isp] += (2.0 * phi - 1.0) / sp.Nref if functional == "reference" \
                else mu / sp.Nref
        Qtot = Qbar - phi * rr / Q
        if isinstance(sp, Point):
            mon = sp.monomer
            e = np.exp(-mon.vol * sys.fields[mon.id])
            g.dw[mon.id] += -mon.vol * e * (phi * rbar[mon.id] / Q + Qtot / n)
            g.dN.append([])
            continue
        cv = _chain_vjp(sys, sp, phi, rbar, Qtot)
        g.dcell += cv.dtheta
        Nref_bar = -phi * rr / sp.Nref
        if active:
            Nref_bar += -phi * (mu - 1.0) / sp.Nref ** 2
        dN = []
        for b in range(sp.nblocks):
            mid = sp.mid_block[b]
            g.dw[mid] += cv.wbar[b]
            ds = sp.ds_block[b]
            dds = cv.dds_ld[b] + float(np.sum(sys.fields[mid] * cv.wbar[b])) / ds
            dds += phi * float(np.sum(rbar[mid] * sp.density_block[b])) / ds
            dNref_b = dds * ds / sp.Nref_block[b] + Nref_bar
            dN.append(sp.vol_block[b] * dNref_b)
        g.dN.append(dN)
    g.dchi = {}
    g.du = {}
    for itx in sys.interactions:
        if hasattr(itx, "chis"):
            for (a, b), _ in sorted(itx.chis.items()):
                va, vb = sys.monomers[a].vol, sys.monomers[b].vol
                g.dchi[(a, b)] = float(np.sum(sys.density[a] * sys.density[b])) / (va * vb * n)
        if hasattr(itx, "uvals"):
            for mid in sorted(itx.mids):
                v = sys.monomers[mid].vol
                g.du[mid] = float(np.sum(0.5 * (sys.density[mid] / v) ** 2)) / n
    return g


def density_vjp(sys, rho_bar, Q_bar):
    """VJP of ``density!(sys)`` alone: given cotangents rho_bar {mid: grid} of the system
    densities and Q_bar (per species) of the species Q, returns (dw {mid: grid},
    dcell).  Runs ``density_update`` first."""
    _check(sys, "reference")
    sys.density_update()
    n = sys.ngrid
    dw = {mid: np.zeros(sys.shape) for mid in sys.mids}
    dcell = np.zeros(sys.cell.nparams)
    for isp, sp in enumerate(sys.species):
        phi, Q = sys.phi_species[isp], sp.Q
        rr = sum(float(np.sum(rho_bar[mid] * sp.density[mid])) for mid in sp.mids)
        Qtot = Q_bar[isp] - phi * rr / Q
        if isinstance(sp, Point):
            mon = sp.monomer
            e = np.exp(-mon.vol * sys.fields[mon.id])
            dw[mon.id] += -mon.vol * e * (phi * rho_bar[mon.id] / Q + Qtot / n)
            continue
        cv = _chain_vjp(sys, sp, phi, rho_bar, Qtot)
        dcell += cv.dtheta
        for b in range(sp.nblocks):
            dw[sp.mid_block[b]] += cv.wbar[b]
    return dw, dcell
