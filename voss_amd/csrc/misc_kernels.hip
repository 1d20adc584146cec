// misc_kernels.hip — gfx950 kernels for the callers around the sieve:
// Goldbach (existence test and partition counts), Sophie Germain counts over
// a device-resident wheel bitmap, and a 64-bit Montgomery Miller-Rabin for
// batched is_prime / next / prev.
//
// Reference behaviour (library/src/):
//   primes/goldbach_upto.cu:25-46   every even n in [4, N] has a partition
//   primes/goldbach.cpp:21-43       #{p1 <= p2 prime, p1 + p2 = n}
//   primes/sophie_germain_upto.cu:22-33  p and 2p+1 both prime
//   core/miller_rabin.cpp:26-51     deterministic MR, bases 2..37
//   primes/next_prev.cu:22-83       odd scan with MR
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>

#include "misc_kernels.h"

namespace voss {
namespace {

__constant__ uint8_t c_res2bit[30] = {8, 0, 8, 8, 8, 8, 8, 1, 8, 8, 8, 2, 8, 3, 8,
                                      8, 8, 4, 8, 5, 8, 8, 8, 6, 8, 8, 8, 8, 8, 7};
constexpr uint32_t kW30Lo = 0x0D0B0701u;
constexpr uint32_t kW30Hi = 0x1D171311u;
__device__ __forceinline__ uint32_t w30(uint32_t o) {
  return __builtin_amdgcn_perm(kW30Hi, kW30Lo, o | 0x0C0C0C00u);
}

// A load that every lane makes from the same address of memory no kernel
// writes (a scalar load through the constant address space).
__device__ __forceinline__ uint32_t load_uniform_u32(const uint32_t* p, uint32_t i) {
  return ((const __attribute__((address_space(4))) uint32_t*)p)[i];
}

// Primality of v from a wheel bitmap whose byte 0 is k = 0 (covers up to limit).
__device__ __forceinline__ bool bm_prime(const uint8_t* bm, uint64_t v) {
  if (v < 7) return v == 2 || v == 3 || v == 5;
  const uint64_t k = v / 30;
  const uint32_t b = c_res2bit[(uint32_t)(v - 30 * k)];
  if (b == 8) return false;
  return (bm[k] >> b) & 1;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// ---------------------------------------------------------------- Montgomery MR
struct Mont {
  uint64_t n, ninv, one, r2;
};

__device__ __forceinline__ uint64_t mont_mul(uint64_t a, uint64_t b, const Mont& M) {
  const uint64_t lo = a * b, hi = __umul64hi(a, b);
  const uint64_t m = lo * M.ninv;
  const uint64_t mh = __umul64hi(m, M.n);
  const uint64_t c1 = lo != 0;
  const uint64_t s1 = hi + mh;
  const bool o1 = s1 < hi;
  const uint64_t s2 = s1 + c1;
  const bool o2 = s2 < s1;
  uint64_t r = s2;
  if (o1 || o2 || r >= M.n) r -= M.n;
  return r;
}

__device__ Mont mont_setup(uint64_t n) {
  Mont M;
  M.n = n;
  uint64_t inv = n;  // n * n == 1 mod 8 for odd n: 3 correct bits
#pragma unroll
  for (int i = 0; i < 5; ++i) inv *= 2 - n * inv;
  M.ninv = 0 - inv;  // -n^-1 mod 2^64
  const uint64_t r = (0 - n) % n;  // 2^64 mod n
  M.one = r;
  uint64_t r2 = r;  // 2^64 * 2^k mod n, doubled 64 times -> 2^128 mod n
  for (int i = 0; i < 64; ++i) {
    const uint64_t t = r2 << 1;
    const bool of = (r2 >> 63) != 0;
    r2 = (of || t >= n) ? t - n : t;
  }
  M.r2 = r2;
  return M;
}

// Deterministic strong-probable-prime bases: {2, 7, 61} decide every n < 2^32
// (Jaeschke 1993); the seven bases below decide every n < 2^64 (Sinclair 2011).
// Primality is a fact of n, so the answers equal those of the reference's
// twelve bases 2..37 (core/miller_rabin.cpp:26-51); fewer bases only make
// the primes, which must pass every base, cheaper.
__constant__ uint32_t c_bases32[3] = {2, 7, 61};
__constant__ uint32_t c_bases64[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};

// 0: composite, 1: prime, 2: odd and free of prime factors <= 37 (needs MR).
__device__ __forceinline__ int mr_prescreen(uint64_t n) {
  if (n < 2) return 0;
  if (n == 2 || n == 3) return 1;
  if ((n & 1) == 0) return 0;
  const uint32_t small[11] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (int i = 0; i < 11; ++i) {
    if (n == small[i]) return 1;
    if (n % small[i] == 0) return 0;
  }
  return n < 41 * 41 ? 1 : 2;
}

// Strong probable-prime test of odd n = d * 2^r + 1 to base a.
__device__ bool mr_strong(const Mont& M, uint64_t d, int r, uint64_t a) {
  if (a >= M.n) a %= M.n;
  if (a == 0) return true;
  const uint64_t minus1 = M.n - M.one;  // -1 in Montgomery form
  uint64_t x = M.one, base = mont_mul(a, M.r2, M);
  for (uint64_t e = d; e; e >>= 1) {
    if (e & 1) x = mont_mul(x, base, M);
    base = mont_mul(base, base, M);
  }
  if (x == M.one || x == minus1) return true;
  for (int k = 0; k < r - 1; ++k) {
    x = mont_mul(x, x, M);
    if (x == minus1) return true;
  }
  return false;
}

__device__ __forceinline__ uint32_t mr_base(uint64_t n, uint32_t i) {
  return n < (1ull << 32) ? (i < 3 ? c_bases32[i] : 0u) : c_bases64[i];
}

// One lane, every base in turn (composites usually stop at base 2).
__device__ bool mr_is_prime(uint64_t n) {
  const int pre = mr_prescreen(n);
  if (pre != 2) return pre == 1;
  uint64_t d = n - 1;
  const int r = __builtin_ctzll(d);
  d >>= r;
  const Mont M = mont_setup(n);
  const uint32_t nb = n < (1ull << 32) ? 3 : 7;
  for (uint32_t i = 0; i < nb; ++i)
    if (!mr_strong(M, d, r, mr_base(n, i))) return false;
  return true;
}

}  // namespace

// ---------------------------------------------------------------- kernels
// One wave per 64 values.  Stage 1: every undecided value takes base 2 on its
// own lane.  Stage 2: the survivors' remaining bases (six each) are spread
// over the wave's lanes, so a wave holding a prime spends about two
// exponentiations instead of seven in sequence.
__global__ void __launch_bounds__(64) is_prime_kernel(const uint64_t* x, uint64_t n, uint8_t* out) {
  __shared__ uint64_t s_val[64];
  __shared__ uint32_t s_comp[64];
  const uint32_t lane = threadIdx.x;
  const uint64_t i = (uint64_t)blockIdx.x * 64 + lane;
  const uint64_t v = i < n ? x[i] : 0;
  int st = mr_prescreen(v);
  if (st == 2) {
    uint64_t d = v - 1;
    const int r = __builtin_ctzll(d);
    d >>= r;
    if (!mr_strong(mont_setup(v), d, r, 2)) st = 0;
  }
  const uint64_t surv = __ballot(st == 2);
  const uint32_t ord = __builtin_amdgcn_mbcnt_hi((uint32_t)(surv >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)surv, 0u));
  const uint32_t m = (uint32_t)__builtin_popcountll(surv);  // wave-uniform
  if (st == 2) s_val[ord] = v;
  s_comp[lane] = 0;
  __syncthreads();
  for (uint32_t t = lane; t < 6 * m; t += 64) {
    const uint32_t k = t / 6, b = t % 6 + 1;
    const uint64_t nv = s_val[k];
    const uint32_t a = mr_base(nv, b);
    if (!a) continue;
    uint64_t d = nv - 1;
    const int r = __builtin_ctzll(d);
    d >>= r;
    if (!mr_strong(mont_setup(nv), d, r, a)) s_comp[k] = 1;
  }
  __syncthreads();
  if (st == 2) st = s_comp[ord] ? 0 : 1;
  if (i < n) out[i] = (uint8_t)st;
}

// next: smallest prime > x (x < 2 -> 2); prev: largest prime < x (x <= 2 -> 0).
__global__ void next_prev_kernel(const uint64_t* x, uint64_t n, uint64_t* out, int dir) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t v = x[i];
  uint64_t res = 0;
  if (dir > 0) {
    if (v < 2) {
      res = 2;
    } else {
      uint64_t c = v + 1;
      if (c == 2) res = 2;
      else {
        if ((c & 1) == 0) ++c;
        // c > v stops at the top of uint64 (no prime above 2^64 - 59): res stays 0
        for (uint64_t steps = 0; steps < 100000000ull && c > v; steps += 2, c += 2)
          if (mr_is_prime(c)) { res = c; break; }
      }
    }
  } else {
    if (v <= 2) res = 0;
    else if (v == 3) res = 2;
    else {
      uint64_t c = v - 1;
      if (c == 2) res = 2;
      else {
        if ((c & 1) == 0) --c;
        res = 2;
        for (uint64_t steps = 0; steps < 100000000ull && c >= 3; steps += 2, c -= 2)
          if (mr_is_prime(c)) { res = c; break; }
      }
    }
  }
  out[i] = res;
}

// The n (even, in [n_lo, n_hi], n_lo >= 6) with no partner among the
// ascending odd primes sp[] (those <= n/2 only) such that n - p is prime
// (goldbach_upto.cu:25-46), appended to ce; the slot counter keeps counting
// past cap, so the caller sees how many there were.
//
// Fifteen n at a time: a thread takes the block of n = 30K + 2m, m = 0..14,
// and tries the partners in ascending order for all fifteen at once.  For
// p = 30a + r, n - p = 30(K - a) + 2m - r: bit o (residue w = W30[o]) of
// byte K - a - b_o, where b_o = [w + r >= 30], decides n = 30K + 2m_o with
// m_o = (w + r - 30 b_o) / 2, and the other seven m give n - p divisible by
// 3 or 5.  So one partner resolves its n through two table lookups, the
// bits of byte K - a and of byte K - a - 1 scattered to their m (a table per
// residue r of p, built in LDS).  A wave stops when every lane's n are
// resolved or its partners exceed n / 2.  Blocks K < 4 (n < 120, where n - p
// may be 3 or 5, which the wheel bitmap does not hold) test one n at a time.
constexpr uint32_t kGcThreads = 256;
constexpr uint32_t kGcR = 10;  // residues of odd primes mod 30: 1, 3, 5, 7, 11, 13, 17, 19, 23, 29

__device__ __forceinline__ uint32_t gc_rindex(uint32_t r) {
  // 1->0 3->1 5->2 7->3 11->4 13->5 17->6 19->7 23->8 29->9
  return r == 1 ? 0u : r == 3 ? 1u : r == 5 ? 2u : r == 7 ? 3u : r == 11 ? 4u : r == 13 ? 5u
       : r == 17 ? 6u : r == 19 ? 7u : r == 23 ? 8u : 9u;
}

__global__ void __launch_bounds__(kGcThreads) goldbach_check_kernel(const uint8_t* bm, const uint32_t* sp, uint32_t nsp,
                                                                   uint64_t n_lo, uint64_t n_hi, uint64_t* ce,
                                                                   unsigned long long* nce, uint64_t cap) {
  // tab[b][ri][x]: the n (bits m = 0..14) that byte value x at K - a - b
  // proves for partners of residue index ri
  __shared__ uint16_t tab[2][kGcR][256];
  constexpr uint32_t kR[kGcR] = {1, 3, 5, 7, 11, 13, 17, 19, 23, 29};
  for (uint32_t e = threadIdx.x; e < 2 * kGcR * 256; e += kGcThreads) {
    const uint32_t x = e & 255, ri = (e >> 8) % kGcR, b = e / (kGcR * 256);
    const uint32_t r = kR[ri];
    uint32_t mask = 0;
    for (uint32_t o = 0; o < 8; ++o) {
      const uint32_t s = w30(o) + r, bo = s >= 30 ? 1u : 0u;
      if (bo == b && ((x >> o) & 1u)) mask |= 1u << ((s - 30 * bo) >> 1);
    }
    tab[b][ri][x] = (uint16_t)mask;
  }
  __syncthreads();
  const uint64_t k_lo = n_lo / 30, k_hi = n_hi / 30;
  const uint64_t stride = (uint64_t)gridDim.x * kGcThreads;
  // the loop bound is the same for every lane of a wave up to its last
  // iteration, where some lanes leave; the ballots below count only the
  // lanes still in it
  for (uint64_t K = k_lo + (uint64_t)blockIdx.x * kGcThreads + threadIdx.x; K <= k_hi; K += stride) {
    // the block's n inside [n_lo, n_hi]
    uint32_t U = 0x7FFFu;
    const uint64_t n0 = 30 * K;
    if (n0 < n_lo) U &= 0x7FFFu << (uint32_t)((n_lo - n0) >> 1);
    if (n0 + 28 > n_hi) U &= 0x7FFFu >> (uint32_t)((n0 + 28 - n_hi) >> 1);
    if (K < 4) {
      // small n: one at a time, with 3 and 5 as partners' partners
      for (uint32_t m = 0; m < 15; ++m) {
        if (!((U >> m) & 1u)) continue;
        const uint64_t n = n0 + 2 * m, half = n / 2;
        for (uint32_t i = 0; i < nsp; ++i) {
          const uint64_t p = sp[i];
          if (p > half) break;
          if (bm_prime(bm, n - p)) { U &= ~(1u << m); break; }
        }
      }
    } else {
      // partners p <= 15K + 14 can serve some n of the block; p serves the
      // n with m >= p - 15K.  Four partners a round: their byte loads are
      // independent, so a round waits for one memory latency.
      const uint64_t plast = 15 * K + 14;
      for (uint32_t i = 0; i < nsp; i += 4) {
        const bool live = U != 0 && (uint64_t)load_uniform_u32(sp, i) <= plast;
        if (!__builtin_amdgcn_ballot_w64(live)) break;
        uint32_t hit = 0;
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u) {
          if (i + u >= nsp) break;  // uniform
          const uint32_t p = load_uniform_u32(sp, i + u);
          if (live && (uint64_t)p <= plast) {
            const uint32_t a = p / 30, ri = gc_rindex(p - 30 * a);
            const uint64_t kb = K - a;  // >= K - (K / 2 + 1) >= 1 for K >= 4
            const uint32_t h = (uint32_t)tab[0][ri][bm[kb]] | (uint32_t)tab[1][ri][bm[kb - 1]];
            const uint64_t d = (uint64_t)p > 15 * K ? (uint64_t)p - 15 * K : 0;
            hit |= h & (0x7FFFu << (uint32_t)d);
          }
        }
        U &= ~hit;
      }
    }
    while (U) {
      const uint32_t m = __builtin_ctz(U);
      U &= U - 1;
      const unsigned long long slot = atomicAdd(nce, 1ull);
      if (slot < cap) ce[slot] = n0 + 2 * m;
    }
  }
}

// Goldbach partition counts as bit-plane correlations of the wheel bitmap.
// For n = 30K + s and p = 30k + W[i] (W = 1, 7, ..., 29), the partner
// q = n - p has residue d = (s - W[i]) mod 30 and byte index K - k - borrow
// (borrow = 1 when s < W[i]).  q is a candidate only when d is a wheel
// residue j, so the count over p in [7, n/2] is
//   sum_i #{k : bit_i(bm[k]) and bit_j(bm[K - borrow_i - k])},
// evaluated four low bytes at a time against the byte-reversed high side.
// blockIdx.y = request; blockIdx.x = chunk of kGbBytesPerBlock low bytes;
// partial counts are added into out (zeroed beforehand).
constexpr uint32_t kGbThreads = 256, kGbIters = 16;
constexpr uint64_t kGbBytesPerBlock = 4ull * kGbThreads * kGbIters;

__global__ void __launch_bounds__(kGbThreads) goldbach_count_kernel(const uint8_t* bm, const uint64_t* ns,
                                                                    uint64_t plim, unsigned long long* out) {
  __shared__ unsigned long long red[kGbThreads / 64];
  const uint64_t n = ns[blockIdx.y];
  // partitions p1 + p2 = n with p1 <= p2 and p1 <= plim
  const uint64_t half = min(n / 2, plim);
  const uint64_t kmax = half / 30;  // last low byte that can hold p <= n/2
  const uint64_t kbeg = (uint64_t)blockIdx.x * kGbBytesPerBlock;
  if (n < 4 || kbeg > kmax) return;  // uniform per workgroup
  unsigned long long c = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // p in {2, 3, 5}: 2 + 2 = 4; 3 and 5 with any prime partner q >= p
    if (n == 4 && half >= 2) c = 1;
    if (n > 4) {
      if (3 <= half && bm_prime(bm, n - 3)) ++c;
      if (5 <= half && bm_prime(bm, n - 5)) ++c;
    }
  }
  const uint64_t K = n / 30;
  const uint32_t sres = (uint32_t)(n - 30 * K);
  // per residue i: partner bit j (8 = none) and borrow
  uint32_t jj[8], bo[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint32_t wi = w30(i);
    const uint32_t b = sres < wi ? 1u : 0u;
    jj[i] = c_res2bit[sres + 30 * b - wi];
    bo[i] = b;
  }
  // residues i admissible in the last byte kmax: 30 kmax + W[i] <= half
  uint32_t last_mask = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (30 * kmax + w30(i) <= half) last_mask |= 1u << i;
  if (n > 4) {
    for (uint32_t it = 0; it < kGbIters; ++it) {
      const uint64_t k0 = kbeg + 4ull * (it * kGbThreads + threadIdx.x);
      if (k0 > kmax) break;
      uint32_t lo = *reinterpret_cast<const uint32_t*>(bm + k0);  // k0 is a multiple of 4
      // clear bytes past kmax and residues past n/2 in byte kmax
#pragma unroll
      for (uint32_t t = 0; t < 4; ++t) {
        const uint64_t k = k0 + t;
        const uint32_t m = k < kmax ? 0xFFu : (k == kmax ? last_mask : 0u);
        lo &= ~(0xFFu << (8 * t)) | (m << (8 * t));
      }
      // high side: hb[u] = bm[K - k0 - u], u = 0..4 (all indices >= K - kmax - 4 may go below 0)
      uint32_t hb[5];
#pragma unroll
      for (uint32_t u = 0; u < 5; ++u) {
        const uint64_t off = k0 + u;
        hb[u] = off <= K ? bm[K - off] : 0u;
      }
      const uint32_t h0 = hb[0] | (hb[1] << 8) | (hb[2] << 16) | (hb[3] << 24);  // byte t = bm[K - k0 - t]
      const uint32_t h1 = hb[1] | (hb[2] << 8) | (hb[3] << 16) | (hb[4] << 24);  // byte t = bm[K - 1 - k0 - t]
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if (jj[i] == 8) continue;
        const uint32_t hi = bo[i] ? h1 : h0;
        c += __builtin_popcount(((lo >> i) & 0x01010101u) & ((hi >> jj[i]) & 0x01010101u));
      }
    }
  }
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t w = 0; w < kGbThreads / 64; ++w) t += red[w];
    if (t) atomicAdd(out + blockIdx.y, t);
  }
}

__global__ void zero_u64_kernel(unsigned long long* p, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Sophie Germain: bitmap primes p in [7, N] with 2p + 1 prime.
__global__ void sophie_kernel(const uint8_t* bm, uint64_t N, unsigned long long* out) {
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long c = 0;
  if (k <= N / 30) {
    uint32_t x = bm[k];
    while (x) {
      const uint32_t o = __builtin_ctz(x);
      x &= x - 1;
      const uint64_t p = 30 * k + w30(o);
      if (p >= 7 && p <= N && bm_prime(bm, 2 * p + 1)) ++c;
    }
  }
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// ---------------------------------------------------------------- launchers
hipError_t launch_is_prime(const uint64_t* x, uint64_t n, uint8_t* out, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(is_prime_kernel, dim3((uint32_t)((n + 63) / 64)), dim3(64), 0, st, x, n, out);
  return hipGetLastError();
}
hipError_t launch_next_prev(const uint64_t* x, uint64_t n, uint64_t* out, int dir, hipStream_t st) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(next_prev_kernel, dim3((uint32_t)((n + 63) / 64)), dim3(64), 0, st, x, n, out, dir);
  return hipGetLastError();
}
hipError_t launch_goldbach_check(const uint8_t* bm, const uint32_t* sp, uint32_t nsp, uint64_t n_lo, uint64_t n_hi,
                                 uint64_t* ce, unsigned long long* nce, uint64_t cap, hipStream_t st) {
  if (n_lo < 6 || n_hi < n_lo || (n_lo & 1) || (n_hi & 1)) return hipErrorInvalidValue;
  const uint64_t nblk = n_hi / 30 - n_lo / 30 + 1;  // blocks of fifteen n
  const uint64_t blocks = std::min<uint64_t>((nblk + kGcThreads - 1) / kGcThreads, 1u << 16);  // the kernel strides over the rest
  hipLaunchKernelGGL(goldbach_check_kernel, dim3((uint32_t)blocks), dim3(kGcThreads), 0, st, bm, sp, nsp, n_lo, n_hi,
                     ce, nce, cap);
  return hipGetLastError();
}
hipError_t launch_goldbach_count(const uint8_t* bm, const uint64_t* ns, uint64_t k, uint64_t n_max, uint64_t plim,
                                 unsigned long long* out, hipStream_t st) {
  if (!k) return hipSuccess;
  if (k > 65535) return hipErrorInvalidValue;  // grid.y limit; callers split larger sets
  hipLaunchKernelGGL(zero_u64_kernel, dim3((uint32_t)((k + 255) / 256)), dim3(256), 0, st, out, k);
  const uint64_t chunks = (n_max / 2 / 30) / kGbBytesPerBlock + 1;
  hipLaunchKernelGGL(goldbach_count_kernel, dim3((uint32_t)chunks, (uint32_t)k), dim3(kGbThreads), 0, st, bm, ns,
                     plim, out);
  return hipGetLastError();
}
hipError_t launch_sophie(const uint8_t* bm, uint64_t N, unsigned long long* out, hipStream_t st) {
  const uint64_t nk = N / 30 + 1;
  hipLaunchKernelGGL(sophie_kernel, dim3((uint32_t)((nk + 255) / 256)), dim3(256), 0, st, bm, N, out);
  return hipGetLastError();
}

}  // namespace voss
