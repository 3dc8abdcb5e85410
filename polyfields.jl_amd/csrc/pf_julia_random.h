// Julia 1.5's seeded normal stream and Dict{Int} iteration order (pf_julia_random.cpp).
#pragma once
#include <cstdint>
#include <vector>

#include "pf_host.h"

namespace pf {

// MersenneTwister(seed) of Julia 1.5: dSFMT-19937 seeded with the 32-bit words of seed
class JuliaMT {
   public:
    explicit JuliaMT(uint64_t seed);
    uint64_t next_raw();   // the next [1, 2) draw as its 64-bit pattern
    double rand();         // rand(rng) for Float64
   private:
    void init_by_array(const std::vector<uint32_t>& key);
    void generate();
    uint64_t st[192][2];   // 191 state words and the lung
    int pos = 0;
};

double julia_randn(JuliaMT& g);                              // randn(rng)
void julia_randn_array(JuliaMT& g, double* out, int64_t n);  // randn!(rng, Vector(n))
uint64_t julia_hash_int(int64_t x);                          // hash(x::Int64)
// iteration order of a Dict{Int} whose keys were inserted in the given order
std::vector<int> julia_dict_order(const std::vector<int>& keys);

}  // namespace pf
