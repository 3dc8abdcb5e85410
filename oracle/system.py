"""FieldSystem physics -- restates src/system/system.jl, src/system/mde.jl,
src/system/fft.jl, src/system/compress.jl, src/interactions/*.jl, the density/stress
methods of src/species/multiblock.jl and src/species/point.jl, and src/fileio.jl:14-23.

All arithmetic is FP64.  Grids are numpy arrays of shape (Nz, Ny, Nx) (x fastest).
Per-monomer dictionaries iterate in ascending monomer id; the reference iterates a Julia
``Dict``, whose order only changes the rounding of sums.
"""
import math

import numpy as np
import scipy.fft as sfft

from .cell import ksize
from .chain import Multiblock, Point

_FFT_WORKERS = 1


def set_fft_workers(n):
    """Thread count for the r2c/c2r transforms (the reference threads only FFTW,
    src/system/fft.jl:24-30)."""
    global _FFT_WORKERS
    _FFT_WORKERS = int(n)


def rfft3(a):
    """Unnormalised r2c along all axes, half spectrum on x (src/system/fft.jl:31)."""
    return sfft.rfftn(a, workers=_FFT_WORKERS)


def irfft3(ak, shape):
    """c2r scaled by 1/n (the ``plan_irfft`` ScaledPlan of src/system/fft.jl:32)."""
    return sfft.irfftn(ak, s=shape, workers=_FFT_WORKERS)


CANONICAL, GRAND = "canonical", "grand"


def _julia_log(x):
    """Julia's ``log`` of a Float64: NaN for NaN, -Inf at 0, +Inf at +Inf, DomainError
    (raised here as ValueError) below 0.  A diverging field state (Q overflowing to Inf
    or turning NaN) then runs on to the NaN field error that scft! reports, as in the
    reference, instead of stopping at Python's ``math.log`` domain check."""
    x = float(x)
    if x != x or x == math.inf:
        return x
    if x == 0.0:
        return -math.inf
    if x < 0.0:
        raise ValueError(f"DomainError with {x}: log will only return a complex result "
                         "if called with a complex argument")
    return math.log(x)


class Compressibility:
    """src/system/compress.jl:13-61."""

    def __init__(self, inv_zeta=0.0):
        self.inv_zeta = float(inv_zeta)
        self.strict = abs(self.inv_zeta - 0.0) <= 0.0   # `inv_zeta ≈ 0.0` (atol 0)
        self.field = None

    def update(self, sys):
        """src/system/compress.jl:37-61."""
        rhosum = np.zeros(sys.shape)
        for mid in sys.mids:
            rhosum += sys.density[mid]
        if self.strict:
            field = rhosum - 1.0
            for mid in sys.mids:
                field += sys.fields[mid] - sys.potentials[mid]
            field /= sys.nmonomers
        elif self.inv_zeta > 0.0:
            field = (1.0 / self.inv_zeta) * (rhosum - 1.0)
        else:
            field = np.zeros(sys.shape)
        self.field = field


class FloryInteraction:
    """src/interactions/flory.jl:9-99."""

    def __init__(self):
        self.mids = set()
        self.chis = {}

    def set_interaction(self, alpha, beta, chi):
        """src/interactions/flory.jl:25-37."""
        alpha = getattr(alpha, "id", alpha)
        beta = getattr(beta, "id", beta)
        assert alpha != beta
        pair = (min(alpha, beta), max(alpha, beta))
        self.mids.add(alpha)
        self.mids.add(beta)
        self.chis[pair] = float(chi)

    def energy(self, sys):
        """src/interactions/flory.jl:39-59."""
        e = 0.0
        for (a, b), chi in sorted(self.chis.items()):
            if a in sys.monomers and b in sys.monomers:
                va, vb = sys.monomers[a].vol, sys.monomers[b].vol
                e += np.sum(chi * (sys.density[a] / va) * (sys.density[b] / vb))
        return e / sys.ngrid

    def energy_bulk(self, sys):
        """src/interactions/flory.jl:61-80."""
        e = 0.0
        for (a, b), chi in sorted(self.chis.items()):
            if a in sys.monomers and b in sys.monomers:
                va, vb = sys.monomers[a].vol, sys.monomers[b].vol
                e += chi * (sys.monomer_fracs[a] / va) * (sys.monomer_fracs[b] / vb)
        return e

    def potential(self, sys, alpha, pot):
        """src/interactions/flory.jl:82-99."""
        for beta in sorted(self.mids):
            pair = (min(alpha, beta), max(alpha, beta))
            if pair not in self.chis:
                continue
            chi = self.chis[pair]
            vb = sys.monomers[beta].vol
            pot += chi * (sys.density[beta] / vb)


class EdwardsInteraction:
    """src/interactions/edwards.jl:9-87 (excluded volume u per monomer)."""

    def __init__(self):
        self.mids = set()
        self.uvals = {}

    def set_interaction(self, mid, u):
        mid = getattr(mid, "id", mid)
        self.mids.add(mid)
        self.uvals[mid] = float(u)

    def energy(self, sys):
        e = 0.0
        for mid in sorted(self.mids):
            if mid in sys.monomers:
                u, v = self.uvals[mid], sys.monomers[mid].vol
                e += np.sum((u / 2) * (sys.density[mid] / v) ** 2)
        return e / sys.ngrid

    def energy_bulk(self, sys):
        e = 0.0
        for mid in sorted(self.mids):
            if mid in sys.monomers:
                u, v = self.uvals[mid], sys.monomers[mid].vol
                e += (u / 2) * (sys.monomer_fracs[mid] / v) ** 2
        return e

    def potential(self, sys, mid, pot):
        if mid in self.mids:
            u, v = self.uvals[mid], sys.monomers[mid].vol
            pot += u * (sys.density[mid] / v)


class FieldSystem:
    """src/system/system.jl:22-74.

    ``dims`` is (Nx, Ny, Nz) in the reference's order.  The reference sorts dims
    descending for the grids but plans its FFT on the unsorted tuple (system.jl:55-58),
    so only descending dims work there; they are required here.
    """

    def __init__(self, dims, cell, monomers=(), ensemble=CANONICAL, compress=None,
                 mde="RK2"):
        dims = tuple(int(d) for d in dims)
        if list(dims) != sorted(dims, reverse=True):
            raise ValueError("dims must be given in descending order (system.jl:55-58)")
        if mde not in ("RK2", "RQM4"):
            raise ValueError(f"Invalid propagator method `{mde}` for PseudoSpectralSolver.")
        self.dims = dims
        self.shape = (dims[2], dims[1], dims[0])
        self.kshape = tuple(reversed(ksize(dims)))
        self.ensemble = ensemble
        self.mde = mde
        self.cell = cell
        cell.setup(dims)
        self.compress = compress if compress is not None else Compressibility(0.0)
        self.monomers = {}
        self.monomer_fracs = {}
        self.fields, self.potentials, self.residuals, self.density = {}, {}, {}, {}
        self.interactions = []
        self.species = []
        self.phi_species = []
        self.mu_species = []
        for mon in monomers:
            self.add_monomer(mon)

    # -- bookkeeping ------------------------------------------------------------------
    @property
    def ngrid(self):
        return int(np.prod(self.dims))

    @property
    def nmonomers(self):
        return len(self.monomers)

    @property
    def mids(self):
        return sorted(self.monomers)

    def add_cell(self, cell):
        """src/system/system.jl:102-106."""
        self.cell = cell
        cell.setup(self.dims)

    def add_monomer(self, mon):
        """src/system/system.jl:113-128."""
        self.monomers[mon.id] = mon
        self.monomer_fracs[mon.id] = 0.0
        for d in (self.fields, self.potentials, self.residuals, self.density):
            d[mon.id] = np.zeros(self.shape)

    def add_species(self, species, val):
        """src/system/system.jl:137-161."""
        for mid in species.mids:
            if mid not in self.monomers:
                raise ValueError(f"Species contains monomer with id = {mid} not present in system.")
        species.Q = 0.0
        if isinstance(species, Multiblock):
            species.q = [np.zeros(self.shape) for _ in range(species.Ns + 1)]
            species.qc = [np.zeros(self.shape) for _ in range(species.Ns + 1)]
            species.density_block = {i: np.zeros(self.shape) for i in range(species.nblocks)}
        species.density = {mid: np.zeros(self.shape) for mid in species.mids}
        self.species.append(species)
        if self.ensemble == CANONICAL:
            if val <= 0.0:
                raise ValueError("Bulk volume fraction must be 0 < phi <= 1.")
            self.phi_species.append(float(val))
            self.mu_species.append(0.0)
        else:
            self.phi_species.append(0.0)
            self.mu_species.append(float(val))

    def add_interaction(self, itx):
        """src/system/system.jl:168-179."""
        for mid in itx.mids:
            if mid not in self.monomers:
                raise ValueError(f"Interaction contains monomer with id = {mid} not present in system.")
        self.interactions.append(itx)

    def validate(self):
        """src/system/system.jl:192-237."""
        if not any(np.mean(w) != 0.0 or np.std(w) != 0.0 for w in self.fields.values()):
            raise ValueError("Fields must be initialized before SCFT simulation.")
        if self.ensemble == CANONICAL and not _approx(sum(self.phi_species), 1.0):
            raise ValueError("Species volume fractions do not sum to unity.")
        self.monomer_fractions()
        charge = sum(self.monomer_fracs[m] * mon.charge / mon.vol for m, mon in self.monomers.items())
        if not _approx(charge, 0.0):
            raise ValueError("Bulk system not charge neutral.")
        for mid in self.monomers:
            if not any(mid in sp.mids for sp in self.species):
                raise ValueError(f"Monomer with ID = {mid} not found in any species.")

    # -- propagator ---------------------------------------------------------------------
    def solver_update(self, omega, vol, b, ds):
        """src/system/mde.jl:43-57 -> (LW1, LW2, LD1, LD2)."""
        lw_coeff = ds / 2.0
        ld_coeff = ds * b ** 2 / 6.0
        LW1 = np.exp(-lw_coeff * vol * omega)
        LW2 = np.exp(-lw_coeff * vol * omega / 2.0)
        LD1 = np.exp(-ld_coeff * self.cell.ksq)
        LD2 = np.exp(-ld_coeff * self.cell.ksq / 2.0)
        return LW1, LW2, LD1, LD2

    def propagate(self, ops, qin):
        """src/system/mde.jl:65-107 (RK2 Strang step; RQM4 Richardson)."""
        LW1, LW2, LD1, LD2 = ops
        if self.mde == "RK2":
            qr = LW1 * qin
            qk = rfft3(qr)
            qk *= LD1
            qr = irfft3(qk, self.shape)
            return LW1 * qr
        qr = LW1 * qin
        qk = rfft3(qr) * LD1
        q1 = LW1 * irfft3(qk, self.shape)
        qr = LW2 * qin
        qk = rfft3(qr) * LD2
        qr = irfft3(qk, self.shape)
        qr *= LW1
        qk = rfft3(qr) * LD2
        qr = irfft3(qk, self.shape)
        q2 = LW2 * qr
        return (4.0 * q2 - q1) / 3.0

    # -- species ------------------------------------------------------------------------
    def species_density(self, sp):
        if isinstance(sp, Point):
            return self._point_density(sp)
        return self._chain_density(sp)

    def _point_density(self, pt):
        """src/species/point.jl:36-53."""
        mon = pt.monomer
        w = np.exp(-mon.vol * self.fields[mon.id])
        pt.Q = np.sum(w) / self.ngrid
        pt.density[mon.id] = w / pt.Q

    def _chain_density(self, ch):
        """src/species/multiblock.jl:125-231."""
        q, qc = ch.q, ch.qc
        q[0][...] = 1.0
        for iblk in range(ch.nblocks):
            mon = self.monomers[ch.mid_block[iblk]]
            ops = self.solver_update(self.fields[mon.id], mon.vol, ch.b_block[iblk],
                                     ch.ds_block[iblk])
            for s in range(ch.begin_block[iblk], ch.begin_block[iblk + 1]):   # 1-based s
                q[s] = self.propagate(ops, q[s - 1])
        if ch.nblocks == 1:                                                   # :153-156
            for s in range(ch.Ns + 1):
                qc[s] = q[ch.Ns - s].copy()
        else:
            qc[ch.Ns][...] = 1.0
            for iblk in range(ch.nblocks - 1, -1, -1):
                mon = self.monomers[ch.mid_block[iblk]]
                ops = self.solver_update(self.fields[mon.id], mon.vol, ch.b_block[iblk],
                                         ch.ds_block[iblk])
                for s in range(ch.begin_block[iblk + 1], ch.begin_block[iblk], -1):
                    qc[s - 2] = self.propagate(ops, qc[s - 1])
        ch.Q = np.sum(q[ch.Ns]) / self.ngrid                                # :175
        for iblk in range(ch.nblocks):                                        # :182-216
            bgn, end = ch.begin_block[iblk], ch.begin_block[iblk + 1]
            rho = q[bgn - 1] * qc[bgn - 1]
            rho += q[end - 1] * qc[end - 1]
            for s in range(bgn + 1, end, 2):
                rho += 4.0 * q[s - 1] * qc[s - 1]
            for s in range(bgn + 2, end - 1, 2):
                rho += 2.0 * q[s - 1] * qc[s - 1]
            rho *= ch.ds_block[iblk] / 3.0
            ch.density_block[iblk] = rho
        for iblk in range(ch.nblocks):                                        # :219-221
            ch.density_block[iblk] /= ch.Q * ch.Nref
        for mid in ch.mids:                                                   # :224-228
            acc = np.zeros(self.shape)
            for iblk in ch.mid_map[mid]:
                acc += ch.density_block[iblk]
            ch.density[mid] = acc

    def species_stress(self, sp, exact=False):
        """src/species/multiblock.jl:233-295 (``exact=False``: the reference formula,
        Re*Re only and every entry but k=0 doubled).  ``exact=True`` keeps the Im*Im term
        and weights the half spectrum by its true multiplicity (1 on the kx=0 and Nyquist
        planes); it is the derivative of the continuum functional, used only to report the
        gap to the reference formula."""
        if isinstance(sp, Point):
            return np.zeros(self.cell.nparams)                       # point.jl:55-58
        cell = self.cell
        dQ = np.zeros(cell.nparams)
        n = self.ngrid
        if exact:
            nx = self.dims[0]
            mult = np.full(self.kshape, 2.0)
            mult[..., 0] = 1.0
            if nx % 2 == 0:
                mult[..., nx // 2] = 1.0
        for iblk in range(sp.nblocks):
            bgn, end = sp.begin_block[iblk], sp.begin_block[iblk + 1]
            b = sp.b_block[iblk]
            ds0 = sp.ds_block[iblk]
            for s in range(bgn, end + 1):
                ds = ds0
                if s != bgn and s != end:
                    ds *= 4.0 if s % 2 == 0 else 2.0
                qk1 = rfft3(sp.q[s - 1])
                qk2 = rfft3(sp.qc[s - 1])
                for k in range(cell.nparams):
                    qtmp = qk1 * cell.dksq[k]
                    if exact:
                        qint = np.sum(mult * (qk2.real * qtmp.real + qk2.imag * qtmp.imag))
                    else:
                        flat2 = qk2.real.ravel()
                        flatt = qtmp.real.ravel()
                        qint = flat2[0] * flatt[0] + np.sum(2 * flat2[1:] * flatt[1:])
                    dQ[k] -= b ** 2 * ds * qint / 6.0 / n ** 2
        dQ /= 3.0
        dQ /= sp.Q * sp.Nref
        return dQ

    # -- system-level physics --------------------------------------------------------------
    def muphi(self):
        """src/system/system.jl:248-259."""
        for isp, sp in enumerate(self.species):
            if self.ensemble == CANONICAL:
                self.mu_species[isp] = _julia_log(self.phi_species[isp] / sp.Q)
            else:
                self.phi_species[isp] = sp.Q * math.exp(self.mu_species[isp])

    def density_update(self):
        """src/system/system.jl:266-289 (``density!(sys)``)."""
        for mid in self.mids:
            self.density[mid] = np.zeros(self.shape)
        for sp in self.species:
            self.species_density(sp)
        self.muphi()
        for isp, sp in enumerate(self.species):
            phi = self.phi_species[isp]
            for mid in sp.mids:
                self.density[mid] += phi * sp.density[mid]

    def potentials_update(self):
        """src/system/system.jl:296-310."""
        for mid in self.mids:
            pot = np.zeros(self.shape)
            for itx in self.interactions:
                itx.potential(self, mid, pot)
            self.potentials[mid] = pot

    def residuals_update(self):
        """src/system/system.jl:317-340 (``residuals!``; fields mean-subtracted in place)."""
        self.density_update()
        self.potentials_update()
        self.compress.update(self)
        eta = self.compress.field
        for mid in self.mids:
            self.residuals[mid] = self.potentials[mid] + eta - self.fields[mid]
        if self.ensemble == CANONICAL:
            for mid in self.mids:
                self.fields[mid] = self.fields[mid] - np.mean(self.fields[mid])
                self.residuals[mid] = self.residuals[mid] - np.mean(self.residuals[mid])

    def uniform_fields(self):
        """src/system/system.jl:348-357."""
        self.potentials_update()
        for mid in self.mids:
            w = self.potentials[mid].copy()
            if self.ensemble == CANONICAL:
                w -= np.mean(w)
            self.fields[mid] = w

    def monomer_fractions(self):
        """src/system/system.jl:364-375."""
        for mid in self.monomer_fracs:
            self.monomer_fracs[mid] = 0.0
        for isp, sp in enumerate(self.species):
            for mid in sp.mids:
                self.monomer_fracs[mid] += self.phi_species[isp] * sp.monomer_fraction(mid)
        return self.monomer_fracs

    def scfstress(self, exact=False):
        """src/system/system.jl:384-395."""
        stress = np.zeros(self.cell.nparams)
        self.muphi()
        for isp, sp in enumerate(self.species):
            stress -= self.phi_species[isp] * self.species_stress(sp, exact)
        return stress

    def free_energy(self):
        """src/system/system.jl:402-433 (``mu = phi`` quirk kept, :406-410)."""
        ftrans = 0.0
        for isp, sp in enumerate(self.species):
            phi = self.phi_species[isp]
            mu = self.phi_species[isp]
            if phi > 1e-8:
                ftrans += phi * (mu - 1.0) / sp.Nref
        ffield = 0.0
        for mid in self.mids:
            ffield -= np.sum(self.density[mid] * self.fields[mid])
        ffield /= self.ngrid
        fexc = sum(itx.energy(self) for itx in self.interactions)
        return ftrans + ffield + fexc

    def free_energy_bulk(self):
        """src/system/system.jl:440-458."""
        f = 0.0
        for isp, sp in enumerate(self.species):
            phi = self.phi_species[isp]
            if phi > 1e-8:
                f += phi * (math.log(phi) - 1.0) / sp.Nref
        self.monomer_fractions()
        for itx in self.interactions:
            f += itx.energy_bulk(self)
        return f

    def field_error(self):
        """src/system/system.jl:470-486."""
        rsum = sum(float(np.sum(self.residuals[m] ** 2)) for m in self.mids)
        wsum = np.finfo(float).eps + sum(float(np.sum(self.fields[m] ** 2)) for m in self.mids)
        return math.sqrt(rsum) / math.sqrt(wsum)


def _approx(a, b):
    """Julia ``≈`` (rtol = sqrt(eps), atol = 0)."""
    return a == b or abs(a - b) <= math.sqrt(np.finfo(float).eps) * max(abs(a), abs(b))


def fieldinit(sys, seed=-1, scale=0.1, fields=None):
    """src/fileio.jl:14-23: ``Random.seed!(seed)``, then ``scale * randn(dims)`` for each
    monomer in the iteration order of the Dict of fields, then ``residuals!``.

    The draws restate Julia 1.5's MersenneTwister and randn (``julia_rng``), so the
    fields equal the reference's for the same seed; a negative seed draws a fresh one.
    ``fields`` (a dict of grids) replaces the random draw, as in the dict form at
    src/fileio.jl:25-54.
    """
    if fields is None:
        from . import julia_rng
        rng = julia_rng.seeded(seed if seed >= 0 else int(np.random.default_rng().integers(2 ** 62)))
        for mid in julia_rng.dict_order(list(sys.monomers)):
            sys.fields[mid] = scale * np.array(julia_rng.randn_array(rng, sys.ngrid)).reshape(sys.shape)
    else:
        for mid, w in fields.items():
            sys.fields[mid] = np.array(w, dtype=float).reshape(sys.shape)
    sys.residuals_update()
