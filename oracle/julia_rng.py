"""Julia 1.5's seeded normal draws, restated -- TEST INFRASTRUCTURE ONLY.

The reference seeds fields with ``Random.seed!(seed)`` and ``scale * randn(sys.dims)``
per monomer in ``Dict`` order (``src/fileio.jl:14-23``), and its known-answer tests
(``test/test_system.jl:22-34``, ``test/test_updaters.jl:17-30``) print values of the
resulting states.  The generator lives in Julia's standard library ``Random`` (Julia 1.5,
the version the reference targets), which is not part of the reference tree.  Its
published algorithms, restated here:

* ``MersenneTwister`` is dSFMT-19937 (Saito and Matsumoto, "SIMD-oriented Fast Mersenne
  Twister", the double-precision variant, MEXP = 19937, N = 191 128-bit words).
  ``seed!(rng, n::Integer)`` initialises it with ``dsfmt_init_by_array`` on the 32-bit
  words of n, least significant first (``make_seed``).  Draws come from the [1, 2)
  doubles in state order; ``rand()`` is such a draw minus 1.
* ``randn`` is the Marsaglia-Tsang ziggurat with 256 layers on 52 random bits: bit 0 of
  the 52 gives the sign, the next 51 bits ``rabs``, ``rabs & 0xFF`` the layer; ``x =
  ±rabs * w[layer]``, accepted at once if ``rabs < k[layer]``, else the base strip (layer
  0: the exponential tail beyond r) or the wedge test against ``f``.  The tables follow
  from r = 3.6541528853610088 and the layer area v = r f(r) + sqrt(pi/2) erfc(r/sqrt 2),
  f(x) = exp(-x^2/2), generated here in 60-digit arithmetic and rounded once.
* ``randn!(::MersenneTwister, A::Array{Float64})`` with 13 or more elements first takes
  length(A) consecutive [1, 2) draws, then turns each into a normal in order (a rejected
  draw pulls its extra numbers from the stream at that point); shorter arrays draw
  element by element.
* ``Dict{Int}`` iterates its 16-slot open-addressing table in slot order (linear probing
  from ``hash(key) & 15``; no rehash below 11 keys), with Julia's
  ``hash(x::Int64) = hash_64_64((3h - |x|) + bits(Float64(x)))``, h = 0 (Thomas Wang's
  64-bit integer mix).

Pinned by the reference's own known answers: the C-monomer draws alone fix the point
species' Q (test/test_system.jl:29), and the same fields then give the chain's Q, the free
energy, the stress and two updater trajectories (tests/test_oracle_julia_rng.py).
"""
import math

import mpmath

_M64 = (1 << 64) - 1
_M32 = (1 << 32) - 1

# dSFMT-19937 parameters (dSFMT-params19937.h)
_N = 191
_POS1 = 117
_SL1 = 19
_SR = 12
_MSK = (0x000ffafffffffb3f, 0x000ffdfffc90fffd)
_FIX = (0x90014964b32f4329, 0x3b8d12ac548a7c7a)
_PCV = (0x3d84e1ac0dc82880, 0x0000000000000001)
_LOW_MASK = 0x000FFFFFFFFFFFFF
_HIGH_CONST = 0x3FF0000000000000


class DSFMT:
    """dSFMT-19937: state words st[i] = (lo, hi) 64-bit pairs, i = 0..N (N: the lung)."""

    def __init__(self, key):
        self.st = [[0, 0] for _ in range(_N + 1)]
        self._init_by_array(list(key))
        self.buf = []
        self.pos = 0

    # 32-bit view: word 2k of a pair is the low half of lo, little-endian order
    def _get32(self, i):
        w, r = divmod(i, 4)
        v = self.st[w][r // 2]
        return (v >> 32) & _M32 if r % 2 else v & _M32

    def _set32(self, i, x):
        w, r = divmod(i, 4)
        v = self.st[w][r // 2]
        x &= _M32
        v = (v & _M32) | (x << 32) if r % 2 else (v & (_M32 << 32)) | x
        self.st[w][r // 2] = v

    def _init_by_array(self, key):
        size = (_N + 1) * 4
        lag = 11 if size >= 623 else (7 if size >= 68 else (5 if size >= 39 else 3))
        mid = (size - lag) // 2
        f1 = lambda x: ((x ^ (x >> 27)) * 1664525) & _M32
        f2 = lambda x: ((x ^ (x >> 27)) * 1566083941) & _M32
        byte = 0x8b8b8b8b
        for w in self.st:
            w[0] = w[1] = (byte << 32) | byte
        g, s = self._get32, self._set32
        count = max(len(key) + 1, size)
        r = f1(g(0) ^ g(mid % size) ^ g((size - 1) % size))
        s(mid % size, g(mid % size) + r)
        r = (r + len(key)) & _M32
        s((mid + lag) % size, g((mid + lag) % size) + r)
        s(0, r)
        count -= 1
        i, j = 1, 0
        while j < count and j < len(key):
            r = f1(g(i) ^ g((i + mid) % size) ^ g((i + size - 1) % size))
            s((i + mid) % size, g((i + mid) % size) + r)
            r = (r + key[j] + i) & _M32
            s((i + mid + lag) % size, g((i + mid + lag) % size) + r)
            s(i, r)
            i = (i + 1) % size
            j += 1
        while j < count:
            r = f1(g(i) ^ g((i + mid) % size) ^ g((i + size - 1) % size))
            s((i + mid) % size, g((i + mid) % size) + r)
            r = (r + i) & _M32
            s((i + mid + lag) % size, g((i + mid + lag) % size) + r)
            s(i, r)
            i = (i + 1) % size
            j += 1
        for _ in range(size):
            r = f2((g(i) + g((i + mid) % size) + g((i + size - 1) % size)) & _M32)
            s((i + mid) % size, g((i + mid) % size) ^ r)
            r = (r - i) & _M32
            s((i + mid + lag) % size, g((i + mid + lag) % size) ^ r)
            s(i, r)
            i = (i + 1) % size
        # initial mask: every state word a double in [1, 2)
        for k in range(_N):
            for h in range(2):
                self.st[k][h] = (self.st[k][h] & _LOW_MASK) | _HIGH_CONST
        # period certification
        t0 = self.st[_N][0] ^ _FIX[0]
        t1 = self.st[_N][1] ^ _FIX[1]
        inner = (t0 & _PCV[0]) ^ (t1 & _PCV[1])
        sh = 32
        while sh > 0:
            inner ^= inner >> sh
            sh >>= 1
        if inner & 1 == 0:
            self.st[_N][1] ^= 1

    def _recursion(self, a, b, lung):
        t0, t1 = a
        l0, l1 = lung
        n0 = ((t0 << _SL1) & _M64) ^ (l1 >> 32) ^ ((l1 << 32) & _M64) ^ b[0]
        n1 = ((t1 << _SL1) & _M64) ^ (l0 >> 32) ^ ((l0 << 32) & _M64) ^ b[1]
        lung[0], lung[1] = n0, n1
        return [(n0 >> _SR) ^ (n0 & _MSK[0]) ^ t0, (n1 >> _SR) ^ (n1 & _MSK[1]) ^ t1]

    def _gen_all(self):
        st = self.st
        lung = list(st[_N])
        for i in range(_N):
            b = st[i + _POS1] if i < _N - _POS1 else st[i + _POS1 - _N]
            st[i] = self._recursion(st[i], b, lung)
        st[_N] = lung
        self.buf = [w for k in range(_N) for w in st[k]]
        self.pos = 0

    def next_raw(self):
        """The next draw as the 64-bit pattern of a double in [1, 2)."""
        if self.pos >= len(self.buf):
            self._gen_all()
        v = self.buf[self.pos]
        self.pos += 1
        return v

    def rand(self):
        """Julia's rand(rng) for Float64: a [1, 2) draw minus 1."""
        return _f64(self.next_raw()) - 1.0


def _f64(bits):
    import struct
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def _ziggurat_tables():
    mpmath.mp.dps = 60
    r = mpmath.mpf("3.6541528853610088")
    f = lambda x: mpmath.exp(-x * x / 2)
    v = r * f(r) + mpmath.sqrt(mpmath.pi / 2) * mpmath.erfc(r / mpmath.sqrt(2))
    m1 = mpmath.mpf(2) ** 51
    xs = {255: r}
    x = r
    for i in range(254, 0, -1):
        x = mpmath.sqrt(-2 * mpmath.log(v / x + f(x)))
        xs[i] = x
    q = v / f(r)
    ki = [0] * 256
    wi = [0.0] * 256
    fi = [0.0] * 256
    ki[0] = int(mpmath.floor(r / q * m1))
    wi[0] = float(q / m1)
    wi[255] = float(r / m1)
    fi[0] = 1.0
    fi[255] = float(f(r))
    for i in range(254, 0, -1):
        ki[i + 1] = int(mpmath.floor(xs[i] / xs[i + 1] * m1))
        fi[i] = float(f(xs[i]))
        wi[i] = float(xs[i] / m1)
    return ki, wi, fi, float(r)


_KI, _WI, _FI, _R = _ziggurat_tables()


def _randn_from(rng, raw):
    r = raw & 0x000fffffffffffff
    rabs = r >> 1
    idx = rabs & 0xFF
    x = (-rabs if r & 1 else rabs) * _WI[idx]
    if rabs < _KI[idx]:
        return x
    return _randn_unlikely(rng, idx, rabs, x)


def _randn_unlikely(rng, idx, rabs, x):
    if idx == 0:
        while True:
            xx = -(1.0 / _R) * math.log(rng.rand())
            yy = -math.log(rng.rand())
            if yy + yy > xx * xx:
                return -_R - xx if (rabs >> 8) & 1 else _R + xx
    if (_FI[idx - 1] - _FI[idx]) * rng.rand() + _FI[idx] < math.exp(-0.5 * x * x):
        return x
    return randn(rng)


def randn(rng):
    """Julia's randn(rng) for Float64."""
    return _randn_from(rng, rng.next_raw())


def randn_array(rng, count, bulk=None):
    """Julia's randn!(rng, Array{Float64}(undef, count)): bulk raw draws for 13 or more
    elements, element by element below that (``bulk`` forces either path)."""
    if bulk is None:
        bulk = count >= 13
    if not bulk:
        return [randn(rng) for _ in range(count)]
    raws = [rng.next_raw() for _ in range(count)]
    return [_randn_from(rng, raw) for raw in raws]


def make_seed(n):
    """Julia 1.5's make_seed(n::Integer): the 32-bit words of n, least significant
    first."""
    if n < 0:
        raise ValueError("seed must be non-negative")
    out = []
    while True:
        out.append(n & _M32)
        n >>= 32
        if n == 0:
            return out


def seeded(seed):
    """Random.seed!(seed) on a MersenneTwister."""
    return DSFMT(make_seed(seed))


def hash_int(x):
    """Julia's hash(x::Int64) with h = 0."""
    a = (0 - abs(x) + _float_bits(float(x))) & _M64
    a = (~a + (a << 21)) & _M64
    a ^= a >> 24
    a = (a + (a << 3) + (a << 8)) & _M64
    a ^= a >> 14
    a = (a + (a << 2) + (a << 4)) & _M64
    a ^= a >> 28
    a = (a + (a << 31)) & _M64
    return a


def dict_order(keys):
    """Iteration order of a Dict{Int} whose keys were inserted in the given order."""
    if len(keys) > 10:
        raise ValueError("Dict order restated for at most 10 keys")
    slot = [None] * 16
    for k in keys:
        i = hash_int(k) & 15
        while slot[i] is not None:
            i = (i + 1) & 15
        slot[i] = k
    return [k for k in slot if k is not None]


def _float_bits(x):
    import struct
    return struct.unpack("<Q", struct.pack("<d", x))[0]
