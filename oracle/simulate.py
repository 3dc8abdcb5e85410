"""SCFT driver -- restates src/simulate.jl:12-148."""
import math
import time
from dataclasses import dataclass


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


def isconverged(ferr, ftol, serr, stol, vcell=False):
    """src/simulate.jl:37-49."""
    fconverged = False
    sconverged = True
    if ferr < ftol:
        fconverged = True
        if vcell and serr > stol:
            sconverged = False
    return fconverged and sconverged


def scft(sys, updater, domain=None, opts=None, log=None):
    """src/simulate.jl:60-148.  ``log`` (a list) receives (step, ferr, fhelm, serr) after
    every iteration so trajectories can be compared point by point."""
    opts = opts or SCFTOptions()
    t0 = time.perf_counter()
    sys.validate()
    sys.residuals_update()
    updater.setup(sys)
    step = 0
    ferr = sys.field_error()
    fhelm = sys.free_energy()
    serr = 0.0
    vcell = domain is not None
    if vcell:
        domain.setup(sys)
        serr = domain.stress_error()
    converged = isconverged(ferr, opts.ftol, serr, opts.stol, vcell)
    while step < opts.nsteps and not converged:
        updater.step()
        if vcell and step % domain.nskip == 0:
            domain.scfstress_update()
            serr = domain.stress_error()
            if serr > opts.stol:
                domain.step(nsteps=domain.nper, tol=opts.stol)
                serr = domain.stress_error()
        ferr = sys.field_error()
        fhelm = sys.free_energy()
        if log is not None:
            log.append((step, ferr, fhelm, serr))
        if math.isnan(ferr):
            raise RuntimeError("Diverging field updater -- breaking.")
        converged = isconverged(ferr, opts.ftol, serr, opts.stol, vcell)
        if converged:
            break
        step += 1
    sys.residuals_update()
    ferr = sys.field_error()
    fhelm = sys.free_energy()
    if vcell:
        serr = domain.stress_error()
    return SCFTResults(time.perf_counter() - t0, step, converged, ferr, serr, fhelm)
