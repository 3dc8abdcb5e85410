// Seeded normal fields as the reference draws them: fieldinit!(sys; seed, scale)
// (src/fileio.jl:14-23) runs Random.seed!(seed) and then scale * randn(sys.dims) for each
// monomer in the iteration order of the Dict{Int} of fields (src/system/system.jl:35,122).
// The generator is Julia 1.5's standard library Random, which is not part of the
// reference tree; its published algorithms are restated here:
//   * MersenneTwister = dSFMT-19937 (Saito and Matsumoto's double-precision SIMD-oriented
//     Fast Mersenne Twister, N = 191 128-bit words), seeded by dsfmt_init_by_array with
//     the 32-bit words of the seed, least significant first (Random.make_seed); draws are
//     the state's [1, 2) doubles in order, rand() = draw - 1.
//   * randn = the 256-layer Marsaglia-Tsang ziggurat on 52 random bits (sign, 51-bit
//     magnitude, 8-bit layer), tables in pf_ziggurat.h; randn!(::MersenneTwister,
//     ::Array{Float64}) of 13 or more elements first takes that many consecutive draws and
//     then converts them in order, shorter arrays draw element by element.
//   * Dict{Int} iteration: slot order of a 16-slot open-addressing table (linear probing
//     from hash(key) & 15, keys inserted in monomer order; no rehash below 11 keys), with
//     Julia's hash(::Int64) = hash_64_64((3h - |x|) + bits(Float64(x))), h = 0.
// Host code only; no GPU call.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "pf_julia_random.h"
#include "pf_ziggurat.h"

namespace pf {

namespace {
constexpr int kN = 191, kPos1 = 117, kSL1 = 19, kSR = 12;
constexpr uint64_t kMsk[2] = {0x000ffafffffffb3fULL, 0x000ffdfffc90fffdULL};
constexpr uint64_t kFix[2] = {0x90014964b32f4329ULL, 0x3b8d12ac548a7c7aULL};
constexpr uint64_t kPcv[2] = {0x3d84e1ac0dc82880ULL, 0x0000000000000001ULL};
constexpr uint64_t kLowMask = 0x000FFFFFFFFFFFFFULL, kHighConst = 0x3FF0000000000000ULL;

inline uint32_t ini1(uint32_t x) { return (x ^ (x >> 27)) * 1664525u; }
inline uint32_t ini2(uint32_t x) { return (x ^ (x >> 27)) * 1566083941u; }

double bits_to_double(uint64_t b) {
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}
uint64_t double_to_bits(double d) {
    uint64_t b;
    std::memcpy(&b, &d, sizeof b);
    return b;
}
}  // namespace

JuliaMT::JuliaMT(uint64_t seed) {
    std::vector<uint32_t> key;
    do {
        key.push_back((uint32_t)(seed & 0xffffffffULL));
        seed >>= 32;
    } while (seed != 0);
    init_by_array(key);
}

void JuliaMT::init_by_array(const std::vector<uint32_t>& key) {
    // the initialisation runs on the state as 32-bit words, each 64-bit half low word
    // first (the little-endian view of the C library)
    const int size = (kN + 1) * 4;
    const int lag = size >= 623 ? 11 : (size >= 68 ? 7 : (size >= 39 ? 5 : 3));
    const int mid = (size - lag) / 2;
    std::vector<uint32_t> s32((size_t)size, 0x8b8b8b8bu);
    auto w32 = [&](int i) -> uint32_t& { return s32[(size_t)i]; };
    const int klen = (int)key.size();
    int count = klen + 1 > size ? klen + 1 : size;
    uint32_t r = ini1(w32(0) ^ w32(mid % size) ^ w32((size - 1) % size));
    w32(mid % size) += r;
    r += (uint32_t)klen;
    w32((mid + lag) % size) += r;
    w32(0) = r;
    --count;
    int i = 1, j = 0;
    for (; j < count && j < klen; ++j) {
        r = ini1(w32(i) ^ w32((i + mid) % size) ^ w32((i + size - 1) % size));
        w32((i + mid) % size) += r;
        r += key[j] + (uint32_t)i;
        w32((i + mid + lag) % size) += r;
        w32(i) = r;
        i = (i + 1) % size;
    }
    for (; j < count; ++j) {
        r = ini1(w32(i) ^ w32((i + mid) % size) ^ w32((i + size - 1) % size));
        w32((i + mid) % size) += r;
        r += (uint32_t)i;
        w32((i + mid + lag) % size) += r;
        w32(i) = r;
        i = (i + 1) % size;
    }
    for (j = 0; j < size; ++j) {
        r = ini2(w32(i) + w32((i + mid) % size) + w32((i + size - 1) % size));
        w32((i + mid) % size) ^= r;
        r -= (uint32_t)i;
        w32((i + mid + lag) % size) ^= r;
        w32(i) = r;
        i = (i + 1) % size;
    }
    for (int k = 0; k <= kN; ++k)
        for (int h = 0; h < 2; ++h)
            st[k][h] = (uint64_t)s32[(size_t)(4 * k + 2 * h)] |
                       ((uint64_t)s32[(size_t)(4 * k + 2 * h + 1)] << 32);
    for (int k = 0; k < kN; ++k)
        for (int h = 0; h < 2; ++h) st[k][h] = (st[k][h] & kLowMask) | kHighConst;
    // period certification
    uint64_t inner = ((st[kN][0] ^ kFix[0]) & kPcv[0]) ^ ((st[kN][1] ^ kFix[1]) & kPcv[1]);
    for (int s = 32; s > 0; s >>= 1) inner ^= inner >> s;
    if ((inner & 1) == 0) st[kN][1] ^= 1;
    pos = 2 * kN;   // empty: the first draw generates
}

void JuliaMT::generate() {
    uint64_t l0 = st[kN][0], l1 = st[kN][1];
    for (int i = 0; i < kN; ++i) {
        const uint64_t* b = st[i + kPos1 < kN ? i + kPos1 : i + kPos1 - kN];
        const uint64_t t0 = st[i][0], t1 = st[i][1];
        const uint64_t n0 = (t0 << kSL1) ^ (l1 >> 32) ^ (l1 << 32) ^ b[0];
        const uint64_t n1 = (t1 << kSL1) ^ (l0 >> 32) ^ (l0 << 32) ^ b[1];
        l0 = n0;
        l1 = n1;
        st[i][0] = (n0 >> kSR) ^ (n0 & kMsk[0]) ^ t0;
        st[i][1] = (n1 >> kSR) ^ (n1 & kMsk[1]) ^ t1;
    }
    st[kN][0] = l0;
    st[kN][1] = l1;
    pos = 0;
}

uint64_t JuliaMT::next_raw() {
    if (pos >= 2 * kN) generate();
    const uint64_t v = st[pos / 2][pos % 2];
    ++pos;
    return v;
}

double JuliaMT::rand() { return bits_to_double(next_raw()) - 1.0; }

namespace {
double randn_unlikely(JuliaMT& g, int idx, int64_t rabs, double x) {
    if (idx == 0) {
        while (true) {
            const double xx = -(1.0 / kZigR) * std::log(g.rand());
            const double yy = -std::log(g.rand());
            if (yy + yy > xx * xx) return ((rabs >> 8) & 1) ? -kZigR - xx : kZigR + xx;
        }
    }
    if ((kZigF[idx - 1] - kZigF[idx]) * g.rand() + kZigF[idx] < std::exp(-0.5 * x * x)) return x;
    return julia_randn(g);
}

double randn_from(JuliaMT& g, uint64_t raw) {
    const uint64_t r = raw & 0x000fffffffffffffULL;
    const int64_t rabs = (int64_t)(r >> 1);
    const int idx = (int)(rabs & 0xFF);
    const double x = (double)((r & 1) ? -rabs : rabs) * kZigW[idx];
    if ((uint64_t)rabs < kZigK[idx]) return x;
    return randn_unlikely(g, idx, rabs, x);
}

uint64_t hash_64_64(uint64_t a) {
    a = ~a + (a << 21);
    a = a ^ (a >> 24);
    a = a + (a << 3) + (a << 8);
    a = a ^ (a >> 14);
    a = a + (a << 2) + (a << 4);
    a = a ^ (a >> 28);
    a = a + (a << 31);
    return a;
}
}  // namespace

double julia_randn(JuliaMT& g) { return randn_from(g, g.next_raw()); }

void julia_randn_array(JuliaMT& g, double* out, int64_t n) {
    if (n < 13) {
        for (int64_t i = 0; i < n; ++i) out[i] = julia_randn(g);
        return;
    }
    std::vector<uint64_t> raw((size_t)n);
    for (auto& r : raw) r = g.next_raw();
    for (int64_t i = 0; i < n; ++i) out[i] = randn_from(g, raw[(size_t)i]);
}

uint64_t julia_hash_int(int64_t x) {
    const uint64_t a = (uint64_t)(x < 0 ? -x : x);
    return hash_64_64((0 - a) + double_to_bits((double)x));
}

std::vector<int> julia_dict_order(const std::vector<int>& keys) {
    if (keys.size() > 10) throw Error(-2, "Dict order restated for at most 10 keys");
    int slot[16];
    for (int& s : slot) s = -1;
    for (size_t k = 0; k < keys.size(); ++k) {
        int i = (int)(julia_hash_int(keys[k]) & 15u);
        while (slot[i] >= 0) i = (i + 1) & 15;
        slot[i] = (int)k;
    }
    std::vector<int> out;
    for (int s : slot)
        if (s >= 0) out.push_back(keys[(size_t)s]);
    return out;
}

}  // namespace pf
