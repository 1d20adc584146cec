// misc_kernels.h — launchers for Goldbach, Sophie Germain and Miller-Rabin kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace voss {

hipError_t launch_is_prime(const uint64_t* x, uint64_t n, uint8_t* out, hipStream_t st);
// dir > 0: next prime; dir < 0: previous prime (0 when none).
hipError_t launch_next_prev(const uint64_t* x, uint64_t n, uint64_t* out, int dir, hipStream_t st);
// bm: wheel bitmap from byte 0 covering [0, n_hi]; sp: ascending odd primes 3, 5, 7, ...;
// appends to ce (first cap slots) the even n in [n_lo, n_hi] with no partner
// among sp; *nce counts them all.
hipError_t launch_goldbach_check(const uint8_t* bm, const uint32_t* sp, uint32_t nsp, uint64_t n_lo, uint64_t n_hi,
                                 uint64_t* ce, unsigned long long* nce, uint64_t cap, hipStream_t st);
// bm covers [0, n_max] with n_max >= every ns[i]; k <= 65535; partitions
// p1 + p2 = n with p1 <= p2 and p1 <= plim (~0: all).  Writes out[i]
// (zeroed by the launch) from a grid of (chunks of n/2, requests).
hipError_t launch_goldbach_count(const uint8_t* bm, const uint64_t* ns, uint64_t k, uint64_t n_max, uint64_t plim,
                                 unsigned long long* out, hipStream_t st);
// bm covers [0, 2N+1]; adds the count of primes p in [7, N] with 2p+1 prime to *out.
hipError_t launch_sophie(const uint8_t* bm, uint64_t N, unsigned long long* out, hipStream_t st);

}  // namespace voss
