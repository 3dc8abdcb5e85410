"""Reference-shaped host API over the HIP library.

Mirrors PolyFields.jl's user-facing types and generics (names without Julia's ``!``):
``UnitCell``, ``Monomer``, ``Multiblock`` / ``Diblock`` / ``Homopolymer`` / ``Point``,
``FloryInteraction`` / ``EdwardsInteraction``, ``Compressibility``, ``FieldSystem``,
``AndersonUpdater`` / ``EulerUpdater`` / ``MomentumUpdater``, ``DomainUpdater``,
``SCFTOptions`` / ``SCFTResults``, ``fieldinit``, ``scft``.  Grids are numpy float64
arrays of shape (Nz, Ny, Nx) -- the reference's x-fastest memory order.

The FieldSystem's grids live on the GPU; reading ``sys.fields[mid]`` or
``chain.density[mid]`` copies them to the host.  Nothing here computes physics: every
operation is a call into the library.
"""
import ctypes as C
import math
import os
import warnings
from dataclasses import dataclass

import numpy as np

from ._lib import SaveFn as _SaveFn, ScftOptions as _CO, ScftResults as _CR, call, dptr, iptr

CANONICAL, GRAND = "canonical", "grand"
_GRID = {"field": 0, "density": 1, "residual": 2, "potential": 3}


class UnitCell:
    """``UnitCell(dim, crystal, params...)`` (src/system/cell.jl:31-54)."""

    def __init__(self, dim, crystal, *params):
        if len(params) == 1 and np.ndim(params[0]) == 1:
            params = tuple(params[0])
        self.dim = int(dim)
        self.crystal = str(crystal)
        self._params = np.array(params, dtype=float)
        self._sys = None
        self._geometry()

    def _geometry(self):
        p = self.params
        R, G, vol = np.zeros(9), np.zeros(9), np.zeros(1)
        dGG = np.zeros(9 * len(p))
        call("pf_cell_geometry", self.dim, self.crystal.encode(), dptr(p), len(p), dptr(R),
             dptr(G), dptr(vol), dptr(dGG))
        self.R, self.G, self.volume = R.reshape(3, 3), G.reshape(3, 3), float(vol[0])
        self.dGG = dGG.reshape(3, 3, len(p))

    @property
    def nparams(self):
        return len(self._params)

    @property
    def params(self):
        if self._sys is not None:
            out = np.zeros(len(self._params))
            call("pf_system_cell", self._sys._h, dptr(out), len(out), 0)
            self._params = out
        return self._params.copy()

    @params.setter
    def params(self, value):
        v = np.array(value, dtype=float)
        if self._sys is not None:
            call("pf_system_cell", self._sys._h, dptr(v), len(v), 1)
        self._params = v
        self._geometry()

    def update(self):
        """``update!(cell)`` -- recomputes the basis and the k-space tables."""
        self.params = self.params


class Monomer:
    """``Monomer(; id, vol, charge)`` (src/system/monomer.jl:7-19)."""

    def __init__(self, id, vol=1.0, charge=0.0):
        self.id, self.vol, self.charge = int(id), float(vol), float(charge)

    def __eq__(self, o):
        return (self.id, self.vol, self.charge) == (o.id, o.vol, o.charge)

    def __hash__(self):
        return hash((self.id, self.vol, self.charge))

    def __lt__(self, o):
        return self.id < o.id


def chain_grid(N_block, Ns):
    """``chain_grid(N_block, Ns)`` (src/species/base.jl:27-56) -> (N, begin, ds)."""
    nb = np.array(N_block, dtype=float)
    begin = np.zeros(len(nb) + 1, dtype=np.int32)
    ds = np.zeros(len(nb))
    N = np.zeros(1)
    call("pf_chain_grid", len(nb), dptr(nb), int(Ns), iptr(begin), dptr(ds), dptr(N))
    return float(N[0]), [int(x) for x in begin], list(ds)


class _SpeciesBase:
    _sys = None
    _idx = None

    @property
    def Q(self):
        if self._sys is None:
            return 0.0
        q = np.zeros(1)
        call("pf_system_species_scalars", self._sys._h, self._idx, dptr(q), None, None)
        return float(q[0])

    @property
    def density(self):
        return _SpeciesDensity(self)


class _SpeciesDensity:
    def __init__(self, sp):
        self.sp = sp

    def __getitem__(self, mid):
        sys = self.sp._sys
        out = np.zeros(sys.shape)
        call("pf_system_get_species_density", sys._h, self.sp._idx, int(mid), dptr(out))
        return out

    def keys(self):
        return list(self.sp.mids)


class _Stack:
    def __init__(self, sp, backward):
        self.sp, self.backward = sp, backward

    def __len__(self):
        return self.sp.Ns + 1

    def __getitem__(self, s):
        sys = self.sp._sys
        if s < 0:
            s += self.sp.Ns + 1
        out = np.zeros(sys.shape)
        call("pf_system_get_propagator", sys._h, self.sp._idx, int(s) + 1, int(self.backward),
             dptr(out))
        return out


class Multiblock(_SpeciesBase):
    """``Multiblock(mon_block, N_block, b_block, Ns)`` (src/species/multiblock.jl:32-65)."""

    kind = "multiblock"

    def __init__(self, mon_block, N_block, b_block, Ns=100):
        assert len(mon_block) == len(N_block) == len(b_block)
        self.monomers = sorted(set(mon_block))
        self.mids = [m.id for m in self.monomers]
        self.mid_block = [m.id for m in mon_block]
        self.N_block = [int(x) for x in N_block]
        self.b_block = [float(x) for x in b_block]
        self.N = sum(self.N_block)
        self.Nref_block = [m.vol * n for m, n in zip(mon_block, N_block)]
        self.Nref, self.begin_block, self.ds_block = chain_grid(self.Nref_block, Ns)
        self.f_block = [x / self.Nref for x in self.Nref_block]
        self.mid_map = {}
        for i, m in enumerate(self.mid_block):
            self.mid_map.setdefault(m, []).append(i)
        if Ns != self.begin_block[-1] - 1:
            warnings.warn("Non-integer chain discretization. Adjusting `ds` and `Ns` appropriately.")
        self.Ns = self.begin_block[-1] - 1
        self._Ns_req = int(Ns)

    @property
    def nblocks(self):
        return len(self.N_block)

    def monomer_fraction(self, mid):
        if mid not in self.mids:
            return 0.0
        return sum(self.f_block[i] for i in self.mid_map[mid])

    @property
    def q(self):
        return _Stack(self, 0)

    @property
    def qc(self):
        return _Stack(self, 1)


def Homopolymer(mon, N, b, Ns=100):
    """src/species/multiblock.jl:67-69."""
    return Multiblock([mon], [N], [b], Ns)


def Diblock(amon, bmon, N, f, b1, b2, Ns=100):
    """src/species/multiblock.jl:71-77."""
    N_block = [int(math.trunc(f * N)), int(math.trunc((1 - f) * N))]
    if sum(N_block) != N:
        raise ValueError(f"Non-integer block lengths for N = {N}, f = {f}.")
    return Multiblock([amon, bmon], N_block, [b1, b2], Ns)


class Point(_SpeciesBase):
    """``Point(mon)`` (src/species/point.jl:10-22)."""

    kind = "point"

    def __init__(self, mon):
        self.monomer = mon
        self.mids = [mon.id]
        self.Nref = mon.vol

    def monomer_fraction(self, mid):
        return 1.0 if mid in self.mids else 0.0


class FloryInteraction:
    """src/interactions/flory.jl:9-37."""

    def __init__(self):
        self.mids = set()
        self.chis = {}
        self._systems = []

    def set_interaction(self, alpha, beta, chi):
        alpha = getattr(alpha, "id", alpha)
        beta = getattr(beta, "id", beta)
        assert alpha != beta
        pair = (min(alpha, beta), max(alpha, beta))
        if pair in self.chis:
            warnings.warn(f"Replacing interaction for monomer pair with IDs = {pair}.")
        self.mids.update(pair)
        self.chis[pair] = float(chi)
        for s in self._systems:
            call("pf_system_set_flory", s._h, pair[0], pair[1], float(chi))


class EdwardsInteraction:
    """src/interactions/edwards.jl:9-32."""

    def __init__(self):
        self.mids = set()
        self.uvals = {}
        self._systems = []

    def set_interaction(self, mid, u):
        mid = getattr(mid, "id", mid)
        self.mids.add(mid)
        self.uvals[mid] = float(u)
        for s in self._systems:
            call("pf_system_set_edwards", s._h, mid, float(u))


def set_interaction(itx, *args):
    """``set_interaction!(itx, args...)`` (src/interactions/base.jl:6-8)."""
    itx.set_interaction(*args)


class Compressibility:
    """src/system/compress.jl:13-24."""

    def __init__(self, inv_zeta=0.0):
        self.inv_zeta = float(inv_zeta)


class _GridView:
    def __init__(self, sys, kind):
        self.sys, self.kind = sys, kind

    def __getitem__(self, mid):
        out = np.zeros(self.sys.shape)
        call("pf_system_get_grid", self.sys._h, _GRID[self.kind], int(mid), dptr(out))
        return out

    def __setitem__(self, mid, value):
        if self.kind != "field":
            raise TypeError("only fields can be written")
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(value, dtype=float), self.sys.shape))
        call("pf_system_set_field", self.sys._h, int(mid), dptr(a))

    def keys(self):
        return self.sys.mids

    def items(self):
        return [(m, self[m]) for m in self.sys.mids]

    def values(self):
        return [self[m] for m in self.sys.mids]


class FieldSystem:
    """``FieldSystem(dims, cell; monomers, ensemble, compress, mde)``
    (src/system/system.jl:51-74) on one GPU.

    ``precision``: "fp64" (every grid and transform in float64) or "fp32" (float32
    propagator and transforms, float64 fields, densities and sums).
    """

    def __init__(self, dims, cell, monomers=(), ensemble=CANONICAL, compress=None, mde="RK2",
                 precision="fp64", device=0):
        dims = tuple(int(d) for d in dims)
        if mde not in ("RK2", "RQM4"):
            raise ValueError(f"Invalid propagator method `{mde}` for PseudoSpectralSolver.")
        self.dims = dims
        self.shape = (dims[2], dims[1], dims[0])
        self.ensemble = ensemble
        self.cell = cell
        self.compress = compress if compress is not None else Compressibility(0.0)
        self.precision = precision
        p = np.array(cell.params, dtype=float)
        h = C.c_void_p()
        call("pf_system_create", C.byref(h), dims[0], dims[1], dims[2], cell.dim,
             cell.crystal.encode(), dptr(p), len(p), 0 if ensemble == CANONICAL else 1,
             self.compress.inv_zeta, {"fp64": 0, "fp32": 1}[precision], int(device))
        self._h = h
        if mde == "RQM4":
            call("pf_system_set_mde", h, 1)
        self.mde = mde
        cell._sys = self
        self.monomers = {}
        self.species = []
        self.phi_species = []
        self.interactions = []
        for m in monomers:
            self.add_monomer(m)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                call("pf_system_destroy", h)
            except Exception:
                pass
            self._h = None

    @property
    def mids(self):
        return sorted(self.monomers)

    @property
    def ngrid(self):
        return int(np.prod(self.dims))

    @property
    def fields(self):
        return _GridView(self, "field")

    @property
    def density(self):
        return _GridView(self, "density")

    @property
    def residuals(self):
        return _GridView(self, "residual")

    @property
    def potentials(self):
        return _GridView(self, "potential")

    def add_cell(self, cell):
        """``add_cell!(sys, cell)`` (src/system/system.jl:102-106): the system takes
        ``cell`` (any crystal system of the grid's dimension) in place of its own."""
        p = np.array(cell.params, dtype=float)
        call("pf_system_add_cell", self._h, int(cell.dim), cell.crystal.encode(), dptr(p),
             len(p))
        if self.cell is not None and self.cell is not cell:
            self.cell._sys = None
        self.cell = cell
        cell._sys = self

    def add_monomer(self, mon):
        call("pf_system_add_monomer", self._h, mon.id, mon.vol, mon.charge)
        self.monomers[mon.id] = mon

    def add_species(self, sp, val):
        """``add_species!(sys, species, phi)`` (src/system/system.jl:137-161)."""
        idx = C.c_int()
        if isinstance(sp, Point):
            call("pf_system_add_point", self._h, sp.mids[0], float(val), C.byref(idx))
        else:
            mids = np.array(sp.mid_block, dtype=np.int32)
            Nb = np.array(sp.N_block, dtype=np.int32)
            bb = np.array(sp.b_block, dtype=float)
            nsu = C.c_int()
            call("pf_system_add_multiblock", self._h, len(mids), iptr(mids), iptr(Nb), dptr(bb),
                 sp._Ns_req, float(val), C.byref(idx), C.byref(nsu))
            assert nsu.value == sp.Ns
        sp._sys, sp._idx = self, idx.value
        self.species.append(sp)
        self.phi_species.append(float(val))

    def add_interaction(self, itx):
        """``add_interaction!(sys, itx)`` (src/system/system.jl:168-179)."""
        for m in itx.mids:
            if m not in self.monomers:
                raise ValueError(f"Interaction contains monomer with id = {m} not present in system.")
        if isinstance(itx, FloryInteraction):
            for (a, b), chi in itx.chis.items():
                call("pf_system_set_flory", self._h, a, b, chi)
        else:
            for m, u in itx.uvals.items():
                call("pf_system_set_edwards", self._h, m, u)
        itx._systems.append(self)
        self.interactions.append(itx)

    # reference generics -----------------------------------------------------------
    def validate(self):
        call("pf_validate", self._h)

    def density_update(self):
        """``density!(sys)``."""
        call("pf_density", self._h)

    def residuals_update(self):
        """``residuals!(sys)``."""
        call("pf_residuals", self._h)

    def field_error(self):
        v = C.c_double()
        call("pf_field_error", self._h, C.byref(v))
        return v.value

    def free_energy(self):
        v = C.c_double()
        call("pf_free_energy", self._h, C.byref(v))
        return v.value

    def free_energy_bulk(self):
        v = C.c_double()
        call("pf_free_energy_bulk", self._h, C.byref(v))
        return v.value

    def scfstress(self):
        out = np.zeros(self.cell.nparams)
        call("pf_scfstress", self._h, dptr(out), len(out))
        return out

    def species_scalars(self, isp):
        q, phi, mu = np.zeros(1), np.zeros(1), np.zeros(1)
        call("pf_system_species_scalars", self._h, isp, dptr(q), dptr(phi), dptr(mu))
        return float(q[0]), float(phi[0]), float(mu[0])

    def synchronize(self):
        call("pf_system_synchronize", self._h)

    def set_precision(self, precision):
        """Switch the propagator between "fp64" and "fp32" (state is kept)."""
        call("pf_system_set_precision", self._h, {"fp64": 0, "fp32": 1}[precision])
        self.precision = precision

    def autotune(self, reps=3):
        """Time the candidate kernel schedules of this grid on this system and keep the
        fastest (process-wide); returns the chosen schedule and its time per density!."""
        sched = np.zeros(4, dtype=np.int32)
        ms = C.c_double()
        call("pf_system_autotune", self._h, int(reps), iptr(sched), C.byref(ms))
        keys = ("slab", "pencil", "planar", "split")
        return {k: int(v) for k, v in zip(keys, sched)}, ms.value

    def set_checkpoint(self, K):
        """Checkpointed propagator sweeps: a checkpoint every K slices (-1: sqrt(Ns+1),
        0: full stacks)."""
        call("pf_system_set_checkpoint", self._h, int(K))

    def timing(self, on):
        call("pf_system_timing", self._h, int(bool(on)))

    def timing_read(self, kind):
        ms, cnt = C.c_double(), C.c_int()
        call("pf_system_timing_read", self._h, int(kind), C.byref(ms), C.byref(cnt))
        return ms.value, cnt.value


def fieldinit(sys, source=None, seed=-1, scale=0.1, fields=None, field_type="omega",
              load_cell=True, interpolate=True):
    """``fieldinit!`` (src/fileio.jl:14-64).  Forms: ``fieldinit(sys, seed=, scale=)`` --
    scale*N(0,1) fields drawn as the reference draws them (Julia 1.5's
    ``Random.seed!(seed)`` and ``randn``, monomers in the reference's Dict order; a
    negative seed draws a fresh one), so a seeded state equals the reference's;
    ``fieldinit(sys, {mid: grid}, field_type=, interpolate=)``;
    ``fieldinit(sys, "file.json", load_cell=, interpolate=)``.  ``fields=`` is the dict
    form without interpolation.  Each form ends with ``residuals!``."""
    from . import fileio
    if isinstance(source, str):
        return fileio.fieldinit_file(sys, source, load_cell=load_cell, interpolate=interpolate)
    if isinstance(source, dict):
        return fileio.fieldinit_dict(sys, source, field_type=field_type, interpolate=interpolate)
    if fields is None:
        s = seed if seed >= 0 else int.from_bytes(os.urandom(8), "little")
        call("pf_system_fieldinit_seeded", sys._h, C.c_uint64(s), float(scale))
        return
    for mid, w in fields.items():
        sys.fields[mid] = np.asarray(w, dtype=float).reshape(sys.shape)
    sys.residuals_update()


class _Updater:
    _h = None
    system = None

    def setup(self, sys):
        call("pf_updater_setup", self._h, sys._h)
        self.system = sys

    def step(self):
        call("pf_updater_step", self._h)

    def __del__(self):
        if self._h:
            try:
                call("pf_updater_destroy", self._h)
            except Exception:
                pass
            self._h = None


class AndersonUpdater(_Updater):
    """``AndersonUpdater(; nhist, iter_euler, lam_euler)`` (src/updaters/anderson.jl:36)."""

    def __init__(self, nhist=50, iter_euler=1, lam_euler=0.01):
        self.nhist, self.iter_euler, self.lam_euler = int(nhist), int(iter_euler), float(lam_euler)
        h = C.c_void_p()
        call("pf_updater_anderson", C.byref(h), self.nhist, self.iter_euler, self.lam_euler)
        self._h = h


class EulerUpdater(_Updater):
    """``EulerUpdater(; method, lam)`` (src/updaters/euler.jl:20-28)."""

    def __init__(self, method="PECE", lam=0.05):
        if method not in ("PEC", "PECE"):
            raise ValueError("Invalid `method` parameter for EulerUpdater. Options include: (:PEC, :PECE)")
        self.method, self.method_id, self.lam = method, 1 if method == "PECE" else 0, float(lam)
        h = C.c_void_p()
        call("pf_updater_euler", C.byref(h), self.method_id, self.lam)
        self._h = h


class MomentumUpdater(_Updater):
    """``MomentumUpdater(; method, lam, gamma, beta1, beta2)`` (src/updaters/momentum.jl:30)."""

    def __init__(self, method="MOM", lam=0.01, gamma=0.9, beta1=0.9, beta2=0.999):
        m = {"MOM": 0, "RMS": 1, "ADAM": 2, "NAG": 3}
        if method not in m:
            raise ValueError("Invalid `method` parameter for MomentumUpdater.")
        self.method, self.method_id = method, m[method]
        self.lam, self.gamma = float(lam), float(gamma)
        self.beta1, self.beta2 = float(beta1), float(beta2)
        h = C.c_void_p()
        call("pf_updater_momentum", C.byref(h), m[method], float(lam), float(gamma),
             float(beta1), float(beta2))
        self._h = h


def updater_with_system(cls, sys, **kw):
    """src/updaters/base.jl:16-20."""
    u = cls(**kw)
    u.setup(sys)
    return u


class DomainUpdater:
    """``DomainUpdater(; nper, nskip, lam, pmax, rlxn)`` (src/updaters/domain.jl:20-24)."""

    def __init__(self, nper=10, nskip=200, lam=0.1, pmax=1.0, rlxn=0.5, fd_step=1e-8):
        self.nper, self.nskip = int(nper), int(nskip)
        self.lam, self.pmax, self.rlxn = float(lam), float(pmax), float(rlxn)
        h = C.c_void_p()
        call("pf_domain_create", C.byref(h), self.nper, self.nskip, self.lam, self.pmax, self.rlxn)
        self._h = h
        self.fd_step = float(fd_step)
        call("pf_domain_set_fd_step", self._h, self.fd_step)
        self.system = None

    def setup(self, sys):
        call("pf_domain_setup", self._h, sys._h)
        self.system = sys

    def step(self, nsteps=1, tol=1e-5):
        call("pf_domain_step", self._h, int(nsteps), float(tol))

    def scfstress_update(self):
        call("pf_domain_stress", self._h, 1, None, 0, None)

    @property
    def stress(self):
        n = self.system.cell.nparams
        out = np.zeros(n)
        call("pf_domain_stress", self._h, 0, dptr(out), n, None)
        return out

    def stress_error(self):
        v = C.c_double()
        call("pf_domain_stress", self._h, 0, None, 0, C.byref(v))
        return v.value

    def __del__(self):
        if self._h:
            try:
                call("pf_domain_destroy", self._h)
            except Exception:
                pass
            self._h = None


@dataclass
class SCFTOptions:
    """src/simulate.jl:12-19."""
    nsteps: int = 100
    nsout: int = 0
    nsave: int = 0
    ftol: float = 1e-5
    stol: float = 1e-4
    savepath: str = ""


@dataclass
class SCFTResults:
    """src/simulate.jl:26-33."""
    time: float
    nsteps: int
    converged: bool
    ferr: float
    serr: float
    fhelm: float


def scft(sys, updater, domain=None, opts=None):
    """``scft!(sys, updater; domain, opts)`` (src/simulate.jl:60-148); the loop runs in
    the library.  With ``opts.savepath`` set, the fields are written to it (JSON,
    ``savefields``) every ``opts.nsave`` steps as the reference does
    (src/simulate.jl:124-128), and once more after the final ``residuals!``: the
    reference's own final write names an option that ``SCFTOptions`` lacks
    (src/simulate.jl:140-142), so that write is the evident intent, not a copy."""
    opts = opts or SCFTOptions()
    o = _CO(opts.nsteps, opts.nsout, opts.nsave, opts.ftol, opts.stol)
    r = _CR()
    dh = domain._h if domain is not None else None
    if opts.savepath:
        from . import fileio
        failure = []

        def save(_user, _step):
            if failure:
                return
            try:
                fileio.savefields(sys, opts.savepath)
            except BaseException as err:   # raised again below, after the loop returns
                failure.append(err)

        cb = _SaveFn(save)
        call("pf_scft_save", sys._h, updater._h, dh, C.byref(o), C.byref(r), cb, None)
        if failure:
            raise failure[0]
    else:
        call("pf_scft", sys._h, updater._h, dh, C.byref(o), C.byref(r))
    updater.system = sys
    if domain is not None:
        domain.system = sys
    return SCFTResults(r.time, r.nsteps, bool(r.converged), r.ferr, r.serr, r.fhelm)


class Batch:
    """Many FieldSystems advanced together on one GPU (no counterpart in the reference;
    it runs the reference's scft! for each system).  Systems must share one precision
    and stay alive while the batch exists."""

    def __init__(self, systems):
        self.systems = list(systems)
        arr = (C.c_void_p * len(self.systems))(*[s._h.value for s in self.systems])
        h = C.c_void_p()
        call("pf_batch_create", C.byref(h), arr, len(self.systems))
        self._h = h

    def density_update(self):
        call("pf_batch_density", self._h)

    def timing(self, on):
        call("pf_batch_timing", self._h, int(bool(on)))

    def timing_read(self, kind):
        ms, cnt = C.c_double(), C.c_int()
        call("pf_batch_timing_read", self._h, int(kind), C.byref(ms), C.byref(cnt))
        return ms.value, cnt.value

    def autotune(self, reps=3):
        """Tune the kernel schedule of every grid size in the batch (process-wide)."""
        out = np.zeros(8 * 16, dtype=np.int32)
        cnt = C.c_int()
        call("pf_batch_autotune", self._h, int(reps), iptr(out), 16, C.byref(cnt))
        keys = ("bytes", "nx", "ny", "nz", "slab", "pencil", "planar", "split")
        return [{k: int(v) for k, v in zip(keys, out[8 * i:8 * i + 8])} for i in range(cnt.value)]

    def use_graph(self, on):
        """Replay the density!/residuals! launch sequence as a captured graph (default)
        or launch it eagerly."""
        call("pf_batch_set_graph", self._h, int(bool(on)))

    def timing_lanes(self, kind):
        """Mean number of propagator lanes in one sampled launch of ``kind``."""
        v = C.c_double()
        call("pf_batch_timing_lanes", self._h, int(kind), C.byref(v))
        return v.value

    def residuals_update(self):
        call("pf_batch_residuals", self._h)

    def sweeps(self):
        """The propagator sweeps of density! alone (q and q+ of every chain), no Q,
        Simpson or densities."""
        call("pf_batch_sweeps", self._h)

    def set_precision(self, precision):
        """Switch every system's propagator between "fp64" and "fp32"."""
        call("pf_batch_set_precision", self._h, {"fp64": 0, "fp32": 1}[precision])
        for s in self.systems:
            s.precision = precision

    def scft(self, nhist=10, iter_euler=1, lam_euler=0.01, nsteps=100, ftol=1e-5, mask=None,
             domain=None, stol=1e-5, updater=None):
        """scft!(sys, AndersonUpdater(; nhist, iter_euler, lam_euler); domain) for every
        system, or for those with a true ``mask`` entry; returns one SCFTResults per system
        (ferr NaN: diverged, or not run).  ``domain``: a DomainUpdater whose settings each
        system's own variable-cell updater takes (the object itself is not set up), or None
        for a fixed cell; ``stol`` as in SCFTOptions.  ``updater``: an AndersonUpdater, a
        MomentumUpdater or an EulerUpdater whose settings every system's updater takes,
        in place of the Anderson arguments."""
        res = (_CR * len(self.systems))()
        m = None
        if mask is not None:
            m = np.ascontiguousarray(np.asarray(mask, dtype=np.int32))
            assert m.shape == (len(self.systems),)
        if isinstance(updater, AndersonUpdater):
            nhist, iter_euler, lam_euler = updater.nhist, updater.iter_euler, updater.lam_euler
        elif isinstance(updater, MomentumUpdater):
            call("pf_batch_scft_momentum", self._h, updater.method_id, updater.lam,
                 updater.gamma, updater.beta1, updater.beta2,
                 domain._h if domain is not None else None, int(nsteps), float(ftol),
                 float(stol), iptr(m) if m is not None else None, res)
            return [SCFTResults(r.time, r.nsteps, bool(r.converged), r.ferr, r.serr, r.fhelm)
                    for r in res]
        elif isinstance(updater, EulerUpdater):
            call("pf_batch_scft_euler", self._h, updater.method_id, updater.lam,
                 domain._h if domain is not None else None, int(nsteps), float(ftol),
                 float(stol), iptr(m) if m is not None else None, res)
            return [SCFTResults(r.time, r.nsteps, bool(r.converged), r.ferr, r.serr, r.fhelm)
                    for r in res]
        elif updater is not None:
            raise TypeError("batches run AndersonUpdater, MomentumUpdater or EulerUpdater")
        if domain is not None:
            call("pf_batch_scft_domain", self._h, int(nhist), int(iter_euler),
                 float(lam_euler), domain._h, int(nsteps), float(ftol), float(stol),
                 iptr(m) if m is not None else None, res)
        else:
            call("pf_batch_scft", self._h, int(nhist), int(iter_euler), float(lam_euler),
                 int(nsteps), float(ftol), iptr(m) if m is not None else None, res)
        return [SCFTResults(r.time, r.nsteps, bool(r.converged), r.ferr, r.serr, r.fhelm)
                for r in res]

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                call("pf_batch_destroy", self._h)
            except Exception:
                pass
            self._h = None


def schedule_reset():
    """Forget every tuned kernel schedule (built-in defaults again)."""
    call("pf_schedule_reset")


def schedule_set(precision, dims, slab=-1, pencil=-1, planar=-1, split=-1):
    """Set the kernel schedule of one grid size by hand (-1: built-in default)."""
    v = np.array([slab, pencil, planar, split], dtype=np.int32)
    call("pf_schedule_set", 4 if precision == "fp32" else 8, int(dims[0]), int(dims[1]),
         int(dims[2]), iptr(v))
