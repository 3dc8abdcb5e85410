"""Unit cell geometry -- restates src/system/basis.jl and src/system/cell.jl (FP64).

Matrices are 3x3 numpy arrays indexed [row, col] exactly like the reference's
``SMat3D``: ``R[:, i]`` is lattice vector a_i and ``G[:, i]`` reciprocal vector b_i
(src/system/cell.jl:20-21).  Derivative arrays are indexed [i, j, k] with k the cell
parameter, as in src/system/cell.jl:22-25.
"""
import math

import numpy as np

CRYSTAL1D = ("lamellar",)                                        # basis.jl:3
CRYSTAL2D = ("square", "rectangular", "hexagonal", "oblique")     # basis.jl:4
CRYSTAL3D = ("cubic", "tetragonal", "orthorhombic", "hexagonal",  # basis.jl:5
             "monoclinic", "triclinic", "pnna", "fddd1")

NPARAMS = {
    (1, "lamellar"): 1,
    (2, "square"): 1, (2, "rectangular"): 2, (2, "hexagonal"): 1, (2, "oblique"): 3,
    (3, "cubic"): 1, (3, "tetragonal"): 2, (3, "orthorhombic"): 3, (3, "hexagonal"): 2,
    (3, "monoclinic"): 4, (3, "triclinic"): 6, (3, "pnna"): 1, (3, "fddd1"): 1,
}


def _isapprox(a, b):
    """Julia ``isapprox`` with default tolerances (rtol = sqrt(eps), atol = 0)."""
    return abs(a - b) <= math.sqrt(np.finfo(float).eps) * max(abs(a), abs(b))


def basis_matrix(a, b, c, alpha=math.pi / 2, beta=math.pi / 2, gamma=math.pi / 2):
    """src/system/basis.jl:80-116."""
    p2 = math.pi / 2
    if _isapprox(alpha, p2) and _isapprox(beta, p2) and _isapprox(gamma, p2):
        return np.diag([float(a), float(b), float(c)])
    cos_alpha = 0.0 if _isapprox(alpha, p2) else math.cos(alpha)
    cos_beta = 0.0 if _isapprox(beta, p2) else math.cos(beta)
    if _isapprox(gamma, p2):
        cos_gamma, sin_gamma = 0.0, 1.0
    else:
        cos_gamma, sin_gamma = math.cos(gamma), math.sin(gamma)
    ax = a
    bx = b * cos_gamma
    by = b * sin_gamma
    cx = c * cos_beta
    cy = c * (cos_alpha - cos_beta * cos_gamma) / sin_gamma
    cz = math.sqrt(c ** 2 - cx ** 2 - cy ** 2)
    # SMat3D(ax,0,0, bx,by,0, cx,cy,cz) is column-major: columns are a1, a2, a3.
    return np.array([[ax, bx, cx], [0.0, by, cy], [0.0, 0.0, cz]], dtype=float)


def make_basis(dim, crystal, params):
    """src/system/basis.jl:11-71."""
    p = [float(x) for x in params]
    pi = math.pi
    crystals = {1: CRYSTAL1D, 2: CRYSTAL2D, 3: CRYSTAL3D}.get(dim, ())
    if crystal not in crystals:
        raise ValueError(f"Unrecognized crystal system `{crystal}` for dim = {dim}.")
    if (dim, crystal) in NPARAMS:
        assert len(p) == NPARAMS[(dim, crystal)]
    if dim == 1:
        return basis_matrix(p[0], 1.0, 1.0, pi / 2, pi / 2, pi / 2)
    if dim == 2:
        if crystal == "square":
            return basis_matrix(p[0], p[0], 1.0, pi / 2, pi / 2, pi / 2)
        if crystal == "rectangular":
            return basis_matrix(p[0], p[1], 1.0, pi / 2, pi / 2, pi / 2)
        if crystal == "hexagonal":
            return basis_matrix(p[0], p[0], 1.0, pi / 2, pi / 2, pi / 3)
        return basis_matrix(p[0], p[1], 1.0, pi / 2, pi / 2, p[2])          # oblique
    if crystal == "cubic":
        return basis_matrix(p[0], p[0], p[0], pi / 2, pi / 2, pi / 2)
    if crystal == "tetragonal":
        return basis_matrix(p[0], p[0], p[1], pi / 2, pi / 2, pi / 2)
    if crystal == "orthorhombic":
        return basis_matrix(p[0], p[1], p[2], pi / 2, pi / 2, pi / 2)
    if crystal == "hexagonal":
        return basis_matrix(p[0], p[0], p[1], pi / 2, pi / 2, pi / 3)
    if crystal == "monoclinic":
        return basis_matrix(p[0], p[1], p[2], pi / 2, p[3], pi / 2)
    if crystal == "triclinic":
        return basis_matrix(*p)
    if crystal == "pnna":
        return basis_matrix(2.0 * p[0], math.sqrt(3.0) * p[0], p[0], pi / 2, pi / 2, pi / 2)
    return basis_matrix(p[0], math.sqrt(3.0) * p[0], 2.0 * math.sqrt(3.0) * p[0],  # fddd1
                        pi / 2, pi / 2, pi / 2)


def basis_dimensions(basis):
    """src/system/basis.jl:123-135."""
    lx, ly, lz = basis[0, 0], basis[1, 1], basis[2, 2]
    xy, xz, yz = basis[0, 1], basis[0, 2], basis[1, 2]
    a = lx
    b = math.sqrt(ly ** 2 + xy ** 2)
    c = math.sqrt(lz ** 2 + xz ** 2 + yz ** 2)
    alpha = math.acos((xy * xz + ly * yz) / (b * c))
    beta = math.acos(xz / c)
    gamma = math.acos(xy / b)
    return np.array([a, b, c, alpha, beta, gamma])


def ksize(dims):
    """src/utils.jl:9 -- half spectrum on the first (x) dimension."""
    return (int(math.floor(dims[0] / 2 + 1)), dims[1], dims[2])


def folded_indices(dims):
    """Integer wave indices (jx, jy, jz) on the half spectrum, shape (Nz, Ny, Nxh) each.

    Restates the index folding of src/system/cell.jl:273-278 (and :303-308): y and z
    are folded to |j| (``iy <= Ny/2 ? iy : Ny-iy``), not to signed j.
    """
    nxh, ny, nz = ksize(dims)
    ix = np.arange(nxh)
    iy = np.arange(ny)
    iz = np.arange(nz)
    jy = np.where(2 * iy <= ny, iy, ny - iy)
    jz = np.where(2 * iz <= nz, iz, nz - iz)
    JZ, JY, JX = np.meshgrid(jz, jy, ix, indexing="ij")
    return JX.astype(float), JY.astype(float), JZ.astype(float)


class UnitCell:
    """src/system/cell.jl:14-54 (constructor runs ``update!``, :31-51)."""

    def __init__(self, dim, crystal, *params):
        if len(params) == 1 and np.ndim(params[0]) == 1:
            params = tuple(params[0])
        self.dim = int(dim)
        self.crystal = str(crystal)
        self.params = np.array(params, dtype=float)
        npr = len(self.params)
        self.dR = np.zeros((3, 3, npr))
        self.dG = np.zeros((3, 3, npr))
        self.dRR = np.zeros((3, 3, npr))
        self.dGG = np.zeros((3, 3, npr))
        self.dims = None
        self.ksq = None
        self.dksq = []
        self.update()

    @property
    def nparams(self):
        return len(self.params)

    def setup(self, dims):
        """src/system/cell.jl:72-86."""
        dim = sum(1 for d in dims if d > 1)
        if dim != self.dim:
            raise ValueError("Number of non-singleton dimensions in System and Cell do not match!")
        self.dims = tuple(dims)
        self.ksq_update()
        self.dksq_update()

    def update(self):
        """src/system/cell.jl:96-109."""
        self.basis()
        self.dbasis()
        self.ddbasis()
        if self.dims is not None:
            self.ksq_update()
            self.dksq_update()

    def basis(self):
        """src/system/cell.jl:113-118."""
        self.R = make_basis(self.dim, self.crystal, self.params)
        self.G = 2 * math.pi * np.linalg.inv(self.R).T
        self.volume = float(np.linalg.det(self.R))

    def dbasis(self):
        """src/system/cell.jl:120-231."""
        dR = np.zeros((3, 3, self.nparams))
        dim, cr = self.dim, self.crystal
        p = self.params
        if dim == 1:
            dR[0, 0, 0] = 1.0
        elif dim == 2:
            if cr == "square":
                dR[0, 0, 0] = 1.0
                dR[1, 1, 0] = 1.0
            elif cr == "rectangular":
                dR[0, 0, 0] = 1.0
                dR[1, 1, 1] = 1.0
            elif cr == "hexagonal":
                dR[0, 0, 0] = 1.0
                dR[0, 1, 0] = 0.5
                dR[1, 1, 0] = math.sqrt(0.75)
            elif cr == "oblique":
                a, b, gamma = p
                dR[0, 0, 0] = 1.0
                dR[0, 1, 1] = math.cos(gamma)
                dR[1, 1, 1] = math.sin(gamma)
                dR[0, 1, 2] = -b * math.sin(gamma)
                dR[1, 1, 2] = b * math.cos(gamma)
        else:
            if cr == "cubic":
                dR[0, 0, 0] = dR[1, 1, 0] = dR[2, 2, 0] = 1.0
            elif cr == "tetragonal":
                dR[0, 0, 0] = dR[1, 1, 0] = 1.0
                dR[2, 2, 1] = 1.0
            elif cr == "orthorhombic":
                dR[0, 0, 0] = 1.0
                dR[1, 1, 1] = 1.0
                dR[2, 2, 2] = 1.0
            elif cr == "hexagonal":
                dR[0, 0, 0] = 1.0
                dR[0, 1, 0] = 0.5
                dR[1, 1, 0] = math.sqrt(0.75)
                dR[2, 2, 1] = 1.0
            elif cr == "monoclinic":
                a, b, c, beta = p
                dR[0, 0, 0] = 1.0
                dR[1, 1, 1] = 1.0
                dR[0, 2, 2] = math.cos(beta)
                dR[2, 2, 2] = math.sin(beta)
                dR[0, 2, 3] = -c * math.sin(beta)
                dR[2, 2, 3] = c * math.cos(beta)
            elif cr == "triclinic":
                a, b, c, alpha, beta, gamma = p
                cy, cz = self.R[1, 2], self.R[2, 2]
                sa, ca = math.sin(alpha), math.cos(alpha)
                sb, cb = math.sin(beta), math.cos(beta)
                sg, cg = math.sin(gamma), math.cos(gamma)
                dR[0, 0, 0] = 1.0
                dR[0, 1, 1] = cg
                dR[1, 1, 1] = sg
                dR[0, 2, 2] = cb
                dR[1, 2, 2] = cy / c
                dR[2, 2, 2] = cz / c
                dR[1, 2, 3] = -c * sa / sg
                dR[2, 2, 3] = c ** 2 * (ca - cb * cg) * sa / sg ** 2 / cz
                dR[0, 2, 4] = -c * sb
                dR[1, 2, 4] = c * sb / math.tan(gamma)
                dR[2, 2, 4] = c ** 2 * (-ca / math.tan(gamma) + cb / sg) * sb / sg / cz
                dR[0, 1, 5] = -b * sg
                dR[1, 1, 5] = b * cg
                dR[1, 2, 5] = c * (-ca / math.tan(gamma) + cb / sg) / sg
                dR[2, 2, 5] = c ** 2 * ((2 + math.cos(2 * alpha) + math.cos(2 * beta))
                                        / math.tan(gamma) - ca * cb * (3 + math.cos(2 * gamma))
                                        / sg) / sg ** 2 / cz / 2
            elif cr == "pnna":
                dR[0, 0, 0] = 2.0
                dR[1, 1, 0] = math.sqrt(3.0)
                dR[2, 2, 0] = 1.0
            elif cr == "fddd1":
                dR[0, 0, 0] = 1.0
                dR[1, 1, 0] = math.sqrt(3.0)
                dR[2, 2, 0] = 2.0 * math.sqrt(3.0)
        # dG_ijk = -(1/2pi) sum_lm G_il dR_mlk G_mj   (cell.jl:215-228)
        G = self.G
        dG = np.zeros((3, 3, self.nparams))
        d = dim
        for k in range(self.nparams):
            for i in range(d):
                for j in range(d):
                    for l in range(d):
                        for m in range(d):
                            dG[i, j, k] -= G[i, l] * dR[m, l, k] * G[m, j]
        dG /= 2 * math.pi
        self.dR = dR
        self.dG = dG

    def ddbasis(self):
        """src/system/cell.jl:233-258."""
        Rt = self.R.T
        G, dR, dG = self.G, self.dR, self.dG
        d = self.dim
        dRR = np.zeros((3, 3, self.nparams))
        dGG = np.zeros((3, 3, self.nparams))
        for k in range(self.nparams):
            for i in range(d):
                for j in range(d):
                    for l in range(d):
                        dRR[i, j, k] += Rt[i, l] * dR[l, j, k] + Rt[j, l] * dR[l, i, k]
                        dGG[i, j, k] += G[l, i] * dG[l, j, k] + G[l, j] * dG[l, i, k]
        self.dRR = dRR
        self.dGG = dGG

    def ksq_update(self):
        """src/system/cell.jl:268-285: ksq = |G * j|^2 with folded j (quirk kept)."""
        JX, JY, JZ = folded_indices(self.dims)
        G = self.G
        kx = G[0, 0] * JX + G[0, 1] * JY + G[0, 2] * JZ
        ky = G[1, 0] * JX + G[1, 1] * JY + G[1, 2] * JZ
        kz = G[2, 0] * JX + G[2, 1] * JY + G[2, 2] * JZ
        self.ksq = kx * kx + ky * ky + kz * kz

    def dksq_update(self):
        """src/system/cell.jl:292-320: dksq_k = sum_{l,m<=dim} j_l j_m dGG[l,m,k]."""
        J = folded_indices(self.dims)
        self.dksq = []
        for k in range(self.nparams):
            acc = np.zeros_like(J[0])
            for l in range(self.dim):
                for m in range(self.dim):
                    acc += J[l] * J[m] * self.dGG[l, m, k]
            self.dksq.append(acc)
