"""Species definitions -- restates src/system/monomer.jl, src/species/base.jl,
src/species/multiblock.jl (constructors) and src/species/point.jl (constructor).

Arrays of per-contour data are 0-based here; the reference's 1-based contour indices
(``begin_block``) are kept 1-based in the attributes so that the Simpson parity rules
(src/species/multiblock.jl:257-263) read the same.
"""
import math
import warnings


class Monomer:
    """src/system/monomer.jl:7-19."""

    def __init__(self, id, vol=1.0, charge=0.0):
        self.id = int(id)
        self.vol = float(vol)
        self.charge = float(charge)

    def __eq__(self, other):            # monomer.jl:23 (hash of id, vol, charge)
        return (self.id, self.vol, self.charge) == (other.id, other.vol, other.charge)

    def __hash__(self):
        return hash((self.id, self.vol, self.charge))

    def __lt__(self, other):             # monomer.jl:24
        return self.id < other.id


def chain_grid(N_block, Ns):
    """src/species/base.jl:27-56.  Returns (N, begin_block[1-based], ds_block)."""
    N = sum(N_block)
    nblocks = len(N_block)
    ds = N / Ns
    begin_block = [0] * (nblocks + 1)
    ds_block = [0.0] * nblocks
    bgn = 1
    begin_block[0] = bgn
    for iblk in range(nblocks):
        Ns_half = int(math.floor(N_block[iblk] / ds / 2.0 + 0.5))
        if Ns_half == 0:
            Ns_half = 1
            ds_block[iblk] = N_block[iblk] / 2.0
        else:
            ds_block[iblk] = N_block[iblk] / Ns_half / 2.0
        bgn = bgn + Ns_half * 2
        begin_block[iblk + 1] = bgn
    return N, begin_block, ds_block


class Multiblock:
    """src/species/multiblock.jl:6-65 (linear block copolymer)."""

    kind = "multiblock"

    def __init__(self, mon_block, N_block, b_block, Ns=100):
        assert len(mon_block) == len(N_block) == len(b_block)
        monomers = sorted(set(mon_block))
        self.monomers = monomers
        self.mids = [m.id for m in monomers]
        self.mid_block = [m.id for m in mon_block]
        self.mid_map = {}
        for iblk, mid in enumerate(self.mid_block):
            self.mid_map.setdefault(mid, []).append(iblk)
        self.N = sum(N_block)
        self.N_block = [int(x) for x in N_block]
        self.vol_block = [m.vol for m in mon_block]
        self.Nref_block = [mon_block[i].vol * N_block[i] for i in range(len(mon_block))]
        self.Nref, self.begin_block, self.ds_block = chain_grid(self.Nref_block, Ns)
        self.f_block = [Nr / self.Nref for Nr in self.Nref_block]
        self.b_block = [float(b) for b in b_block]
        if Ns != self.begin_block[-1] - 1:              # multiblock.jl:55-59
            warnings.warn("Non-integer chain discretization. Adjusting `ds` and `Ns` appropriately.")
            Ns = self.begin_block[-1] - 1
        self.Ns = Ns
        self.Q = 0.0
        self.q = None
        self.qc = None
        self.density = {}
        self.density_block = {}

    @property
    def nblocks(self):
        return len(self.N_block)

    def monomer_fraction(self, mid):
        """src/species/multiblock.jl:93-102."""
        if mid not in self.mids:
            return 0.0
        return sum(self.f_block[i] for i in self.mid_map[mid])


def Homopolymer(mon, N, b, Ns=100):
    """src/species/multiblock.jl:67-69."""
    return Multiblock([mon], [N], [b], Ns)


def Diblock(amon, bmon, N, f, b1, b2, Ns=100):
    """src/species/multiblock.jl:71-77 (``trunc`` check kept)."""
    N_block = [int(math.trunc(f * N)), int(math.trunc((1 - f) * N))]
    if sum(N_block) != N:
        raise ValueError(f"Non-integer block lengths for N = {N}, f = {f}.")
    return Multiblock([amon, bmon], N_block, [b1, b2], Ns)


class Point:
    """src/species/point.jl:10-22 (point particle / solvent)."""

    kind = "point"

    def __init__(self, mon):
        self.monomer = mon
        self.mids = [mon.id]
        self.Nref = mon.vol
        self.Q = 0.0
        self.density = {}

    def monomer_fraction(self, mid):
        """src/species/point.jl:28."""
        return 1.0 if mid in self.mids else 0.0
