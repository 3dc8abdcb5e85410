"""Field and cell updaters -- restates src/updaters/anderson.jl, src/updaters/domain.jl,
src/updaters/euler.jl and src/updaters/momentum.jl (FP64)."""
import math

import numpy as np


def lu_inverse(U):
    """Inverse of a small dense matrix by LU with partial pivoting (first maximal
    pivot), then forward/back substitution against the identity.

    Stands in for LAPACK getrf/getri behind ``inv!(lu!(U))`` (src/updaters/anderson.jl:
    184-189) and ``inv(dstress)`` (src/updaters/domain.jl:67).  The C++ host code of the
    product runs the same algorithm in the same order, so both sides round alike.  Raises
    ``ZeroDivisionError`` on an exactly zero pivot, as ``lu!`` raises SingularException.
    """
    n = U.shape[0]
    A = [list(map(float, row)) for row in U]
    perm = list(range(n))
    for k in range(n):
        p = max(range(k, n), key=lambda i: (abs(A[i][k]), -i))
        if A[p][k] == 0.0:
            raise ZeroDivisionError(f"singular matrix: zero pivot in column {k + 1}")
        if p != k:
            A[k], A[p] = A[p], A[k]
            perm[k], perm[p] = perm[p], perm[k]
        for i in range(k + 1, n):
            A[i][k] = A[i][k] / A[k][k]
            f = A[i][k]
            for j in range(k + 1, n):
                A[i][j] = A[i][j] - f * A[k][j]
    inv = np.zeros((n, n))
    for c in range(n):
        y = [0.0] * n
        for i in range(n):
            acc = 1.0 if perm[i] == c else 0.0
            for j in range(i):
                acc -= A[i][j] * y[j]
            y[i] = acc
        x = [0.0] * n
        for i in range(n - 1, -1, -1):
            acc = y[i]
            for j in range(i + 1, n):
                acc -= A[i][j] * x[j]
            x[i] = acc / A[i][i]
        inv[:, c] = x
    return inv


def matvec(M, v):
    """Row-by-row dot in index order (the host code uses the same order)."""
    out = np.zeros(M.shape[0])
    for i in range(M.shape[0]):
        acc = 0.0
        for j in range(M.shape[1]):
            acc += M[i, j] * v[j]
        out[i] = acc
    return out


class AndersonUpdater:
    """src/updaters/anderson.jl:13-228 (Matsen's Anderson mixing), history kept as the
    reference keeps it: ``nhist+1`` slots with a full copy-shift once full."""

    def __init__(self, nhist=50, iter_euler=1, lam_euler=0.01):
        self.nhist = int(nhist)
        self.iter_euler = max(1, int(iter_euler))            # anderson.jl:38
        self.lam_euler = float(lam_euler)
        self.iter = 0
        self.iter_hist = 0
        self.lam = 0.0
        self.is_setup = False
        self.system = None

    def setup(self, sys):
        """src/updaters/anderson.jl:54-94."""
        self.system = sys
        if self.is_setup:
            return
        if sys.ensemble != "canonical":
            raise ValueError("Anderson mixing can only be used for systems in the canonical ensemble.")
        self.res_hist = [{m: np.zeros(sys.shape) for m in sys.mids} for _ in range(self.nhist + 1)]
        self.field_hist = [{m: np.zeros(sys.shape) for m in sys.mids} for _ in range(self.nhist + 1)]
        for m in sys.mids:
            self.field_hist[0][m] = sys.fields[m].copy()
        self.iter = 0
        self.iter_hist = 0
        self.lam = 0.0
        self.is_setup = True

    def step(self):
        """src/updaters/anderson.jl:98-228."""
        sys = self.system
        self.iter += 1
        if self.iter < self.nhist + 1:
            self.iter_hist = self.iter - 1
            self.lam = 1.0 - 0.9 ** self.iter
        else:
            self.iter_hist = self.nhist
            self.lam = 1.0
        it, ih = self.iter, self.iter_hist
        sys.residuals_update()
        for m in sys.mids:
            self.res_hist[ih][m] = sys.residuals[m].copy()
        if it <= self.iter_euler:
            for m in sys.mids:
                sys.fields[m] = sys.fields[m] + self.lam_euler * sys.residuals[m]
        else:
            U = np.zeros((ih, ih))
            V = np.zeros(ih)
            res_iter = self.res_hist[ih]
            field_iter = self.field_hist[ih]
            for mm in range(1, ih + 1):
                res_mm = self.res_hist[ih - mm]
                for nn in range(mm, ih + 1):
                    res_nn = self.res_hist[ih - nn]
                    for m in sys.mids:
                        r = res_iter[m]
                        U[mm - 1, nn - 1] += float(np.sum((r - res_mm[m]) * (r - res_nn[m])))
                    U[nn - 1, mm - 1] = U[mm - 1, nn - 1]
                for m in sys.mids:
                    r = res_iter[m]
                    V[mm - 1] += float(np.sum((r - res_mm[m]) * r))
            try:
                C = matvec(lu_inverse(U), V)
            except ZeroDivisionError as err:
                raise RuntimeError(f"Error inverting Umn matrix on iteration {it}: {err}")
            self.U, self.V, self.C = U, V, C
            for m in sys.mids:
                WW = sys.fields[m].copy()
                DD = res_iter[m].copy()
                for nn in range(1, ih + 1):
                    WW += C[nn - 1] * (self.field_hist[ih - nn][m] - field_iter[m])
                    DD += C[nn - 1] * (self.res_hist[ih - nn][m] - res_iter[m])
                sys.fields[m] = WW + self.lam * DD
        if it < self.nhist + 1:
            for m in sys.mids:
                self.field_hist[it][m] = sys.fields[m].copy()
        else:
            for nh in range(self.nhist):
                for m in sys.mids:
                    self.res_hist[nh][m] = self.res_hist[nh + 1][m].copy()
                    self.field_hist[nh][m] = self.field_hist[nh + 1][m].copy()
            for m in sys.mids:
                self.field_hist[self.nhist][m] = sys.fields[m].copy()


class DomainUpdater:
    """src/updaters/domain.jl:7-120 (variable cell, forward-difference Jacobian)."""

    def __init__(self, nper=10, nskip=200, lam=0.1, pmax=1.0, rlxn=0.5, fd_step=1e-8):
        self.nper, self.nskip = int(nper), int(nskip)
        self.lam, self.pmax, self.rlxn = float(lam), float(pmax), float(rlxn)
        self.fd_step = float(fd_step)   # 1e-8 in the reference (domain.jl:101)
        self.system = None

    def stress_error(self):
        """src/updaters/domain.jl:30."""
        return float(np.linalg.norm(self.stress))

    def setup(self, sys):
        """src/updaters/domain.jl:32-42."""
        self.system = sys
        npr = sys.cell.nparams
        self.stress = sys.scfstress()
        self.oldparams = np.zeros(npr)
        self.step_vec = np.zeros(npr)
        self.dstress = np.zeros((npr, npr))

    def scfstress_update(self):
        """src/updaters/domain.jl:44-48."""
        self.stress = self.system.scfstress()

    def step(self, nsteps=1, tol=1e-5):
        """src/updaters/domain.jl:50-92: exactly ``nsteps`` Newton iterations."""
        sys = self.system
        cell = sys.cell
        self.scfstress_update()
        it = 0
        while it < nsteps:
            self.dstress_update()
            if cell.dim == 1:
                self.step_vec = np.full(cell.nparams, -1.0 * self.stress[0] / self.dstress[0, 0])
            else:
                self.step_vec = -1.0 * matvec(lu_inverse(self.dstress), self.stress)
            size = float(np.linalg.norm(self.step_vec))
            if size > self.pmax:
                self.step_vec = self.step_vec * (self.pmax / size)
            cell.params = cell.params + self.lam * self.step_vec
            cell.params = self.rlxn * cell.params + (1 - self.rlxn) * self.oldparams
            cell.update()
            sys.density_update()
            sys.residuals_update()
            self.scfstress_update()
            it += 1

    def dstress_update(self):
        """src/updaters/domain.jl:94-120 (dx = 1e-8; params restored without
        ``update!(cell)`` before the last ``density!``)."""
        sys = self.system
        cell = sys.cell
        self.oldparams = cell.params.copy()
        dx = self.fd_step
        x = 1.0 / dx
        for k in range(cell.nparams):
            cell.params = self.oldparams.copy()
            cell.params[k] += dx
            cell.update()
            sys.density_update()
            pstress = sys.scfstress()
            self.dstress[:, k] = x * (pstress - self.stress)
        cell.params = self.oldparams.copy()
        sys.density_update()


class EulerUpdater:
    """src/updaters/euler.jl:12-66 (PEC / PECE)."""

    def __init__(self, method="PECE", lam=0.05):
        if method not in ("PEC", "PECE"):
            raise ValueError("Invalid `method` parameter for EulerUpdater. Options include: (:PEC, :PECE)")
        self.method, self.lam = method, float(lam)

    def setup(self, sys):
        self.system = sys
        self.fields = {m: sys.fields[m].copy() for m in sys.mids}

    def step(self):
        sys = self.system
        for m in sys.mids:
            sys.fields[m] = sys.fields[m] + self.lam * sys.residuals[m]
            self.fields[m] = sys.fields[m] + 0.5 * self.lam * sys.residuals[m]
        sys.residuals_update()
        for m in sys.mids:
            self.fields[m] = self.fields[m] + 0.5 * self.lam * sys.residuals[m]
        for m in sys.mids:
            sys.fields[m] = self.fields[m].copy()
        if self.method == "PECE":
            sys.residuals_update()


class MomentumUpdater:
    """src/updaters/momentum.jl:16-165 (MOM / RMS / ADAM / NAG)."""

    def __init__(self, method="MOM", lam=0.01, gamma=0.9, beta1=0.9, beta2=0.999):
        assert 0.0 < gamma < 1.0 and 0.0 < beta1 < 1.0 and 0.0 < beta2 < 1.0
        if method not in ("MOM", "RMS", "ADAM", "NAG"):
            raise ValueError("Invalid `method` parameter for MomentumUpdater.")
        self.method, self.lam = method, float(lam)
        self.gamma, self.beta1, self.beta2 = float(gamma), float(beta1), float(beta2)

    def setup(self, sys):
        """src/updaters/momentum.jl:47-61."""
        self.system = sys
        self.step_ = {m: sys.residuals[m].copy() for m in sys.mids}
        self.temp = {m: sys.fields[m].copy() for m in sys.mids}
        self.mt = {m: np.zeros(sys.shape) for m in sys.mids}
        self.vt = {m: np.zeros(sys.shape) for m in sys.mids}

    def step(self):
        """src/updaters/momentum.jl:63-84."""
        sys = self.system
        b1, b2 = self.beta1, self.beta2
        if self.method == "NAG":                               # :140-165
            self.temp = {m: sys.fields[m].copy() for m in sys.mids}
            for m in sys.mids:
                sys.fields[m] = sys.fields[m] + self.lam * self.mt[m]
            sys.residuals_update()
            for m in sys.mids:
                self.mt[m] = self.gamma * self.mt[m] + (1 - self.gamma) * sys.residuals[m]
            self.step_ = {m: self.mt[m].copy() for m in sys.mids}
            for m in sys.mids:
                sys.fields[m] = self.temp[m].copy()
        else:
            sys.residuals_update()
            r = sys.residuals
            if self.method in ("MOM", "ADAM"):
                for m in sys.mids:
                    self.mt[m] = b1 * self.mt[m] + (1 - b1) * r[m]
            if self.method in ("RMS", "ADAM"):
                for m in sys.mids:
                    self.vt[m] = b2 * self.vt[m] + (1 - b2) * r[m] ** 2
            for m in sys.mids:
                if self.method == "MOM":
                    self.step_[m] = self.mt[m].copy()
                elif self.method == "RMS":
                    self.step_[m] = r[m] / (np.sqrt(self.vt[m]) + 1e-8)
                else:
                    self.step_[m] = self.mt[m] / (np.sqrt(self.vt[m]) + 1e-8)
        for m in sys.mids:
            sys.fields[m] = sys.fields[m] + self.lam * self.step_[m]


def updater_with_system(cls, sys, **kwargs):
    """src/updaters/base.jl:16-20."""
    u = cls(**kwargs)
    u.setup(sys)
    return u
