// host_math.h — single-number arithmetic on the host CPU, as in the reference
// (library/src/core/miller_rabin.cpp, primes/factorize.cpp,
// primes/mersenne_fermat.cpp, numbers/*.cpp).  Batched forms run on the GPU
// (misc_kernels.hip).
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

namespace voss {

bool host_is_prime(uint64_t n);                  // deterministic MR, bases 2..37
bool host_next_prime(uint64_t x, uint64_t* out); // smallest prime > x
bool host_prev_prime(uint64_t x, uint64_t* out); // largest prime < x (x > 2)
std::vector<uint64_t> host_factorize(uint64_t x);  // ascending, with multiplicity
std::vector<std::pair<uint64_t, uint32_t>> host_factor_exponents(uint64_t x);
bool host_lucas_lehmer(uint32_t p);              // 2^p - 1 prime, 2 <= p <= 63

}  // namespace voss
