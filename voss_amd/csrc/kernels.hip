// kernels.hip — gfx950 kernels of the wheel-30 segmented sieve.
//
// One workgroup sieves one segment (a run of bitmap bytes, one byte per 30
// integers) held entirely in LDS, then walks it in order to produce counts,
// Chebyshev tallies, the consecutive-prime gap histogram, the segment's first
// and last prime and its largest gap.  A second pass (stitch_kernel) joins
// segments: it adds each boundary gap (first prime of a segment minus the last
// prime before it) and flags the few segments that hold maximal-gap records,
// which rescan_kernel then lists in order.
//
// Reference behaviour these kernels reproduce (library/src/primes/):
//   sieve.cu:10-37        crossing-off (here: LDS ds_and_b32, offsets per lane)
//   extract.cu:10-39      per-segment prime count (here: popcounts)
//   gaps.cu:31-75         gap histogram + large gaps (here: in-order LDS walk)
//   chebyshev.cu:11-25    mod-4 tallies (here: per-bit-plane masked popcounts)
//   context.cu:462-501    boundary stitching (here: stitch_kernel)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "common.h"
#include "kernels.h"

namespace voss {
namespace {

constexpr uint32_t kW30Lo = 0x0D0B0701u;  // W30[0..3] = 1, 7, 11, 13
constexpr uint32_t kW30Hi = 0x1D171311u;  // W30[4..7] = 17, 19, 23, 29

// 8 * W30[o] as byte tables for v_perm_b32, in order and reversed.
constexpr uint32_t kW30x8Lo = 0x68583808u;     // 8, 56, 88, 104
constexpr uint32_t kW30x8Hi = 0xE8B89888u;     // 136, 152, 184, 232
constexpr uint32_t kW30x8RevLo = 0x8898B8E8u;  // 232, 184, 152, 136
constexpr uint32_t kW30x8RevHi = 0x08385868u;  // 104, 88, 56, 8

// W30[o] for o in 0..7 (one v_perm_b32: selector bytes 1..3 = 0x0C give 0).
__device__ __forceinline__ uint32_t w30(uint32_t o) {
  return __builtin_amdgcn_perm(kW30Hi, kW30Lo, o | 0x0C0C0C00u);
}

__device__ __forceinline__ uint32_t w30_const(int o) {
  return o == 0 ? 1 : o == 1 ? 7 : o == 2 ? 11 : o == 3 ? 13 : o == 4 ? 17 : o == 5 ? 19 : o == 6 ? 23 : 29;
}

// Index of the lowest set bit of byte k of w, 0xFFFFFFFF for a zero byte:
// one v_ffbl_b32 whose SDWA source select zero-extends the byte (k is a
// constant after unrolling, so the switch folds).
__device__ __forceinline__ uint32_t ffbl_byte(uint32_t w, int k) {
  uint32_t t;
  switch (k & 3) {
    case 0: asm("v_ffbl_b32_sdwa %0, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_0" : "=v"(t) : "v"(w)); break;
    case 1: asm("v_ffbl_b32_sdwa %0, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1" : "=v"(t) : "v"(w)); break;
    case 2: asm("v_ffbl_b32_sdwa %0, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_2" : "=v"(t) : "v"(w)); break;
    default: asm("v_ffbl_b32_sdwa %0, %1 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_3" : "=v"(t) : "v"(w)); break;
  }
  return t;
}

// LDS atomic add without return (ds_add_u32).
__device__ __forceinline__ void lds_add(uint32_t* p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) u32x4 lds_u32x4;
__device__ __forceinline__ void lds_add(lds_u32* p, uint32_t v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// k mod p for k < 2^52, p < 2^24, from a Newton-refined FP64 reciprocal.
// k mod p for k < 2^53, 0 < p < 2^24.  The quotient estimate from a refined
// FP64 reciprocal is within one of the truth, so k - q*p lies in (-p, 2p): an
// integer far below 2^53, which one fused multiply-add yields exactly.
__device__ __forceinline__ uint32_t mod_u64(uint64_t k, uint32_t p) {
  const double pd = (double)p;
  double inv = __builtin_amdgcn_rcp(pd);
  inv = __builtin_fma(inv, __builtin_fma(-pd, inv, 1.0), inv);
  const double kd = (double)k;
  const double qd = __builtin_floor(kd * inv);
  int32_t r = (int32_t)__builtin_fma(-qd, pd, kd);
  if (r < 0) r += (int32_t)p;
  if (r >= (int32_t)p) r -= (int32_t)p;
  return (uint32_t)r;
}

__device__ __forceinline__ uint64_t isqrt_u64(uint64_t n) {
  uint64_t r = (uint64_t)__builtin_sqrt((double)n);
  while (r * r > n) --r;
  while ((r + 1) * (r + 1) <= n) ++r;
  return r;
}

// Number of wheel residues W30[o] that are < r (r in 0..30).
__device__ __forceinline__ uint32_t residues_below(uint32_t r) {
  uint32_t c = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) c += (w30_const(o) < r) ? 1u : 0u;
  return c;
}

// Bit index (within a segment) of the first candidate >= v, clamped to [0, 8*nbytes].
__device__ __forceinline__ uint32_t bit_at_or_above(uint64_t v, uint64_t k0, uint32_t nbytes) {
  if (v <= 30 * k0) return 0;
  uint64_t k = v / 30;
  if (k >= k0 + nbytes) return 8 * nbytes;
  return 8 * (uint32_t)(k - k0) + residues_below((uint32_t)(v % 30));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
  return v;
}

// ---------------------------------------------------------------- LDS layout
// u32 word offsets into dynamic LDS.
// The tables come first: at LDS address 0 a row's address is its index
// times 8, within reach of ds_read2_b64's 8-bit offset fields.
constexpr uint32_t kLdsTab = 0;                                       // 3 x 64 bytes = 48 words
constexpr uint32_t kLdsHist = kLdsTab + 48;                           // kHistBins*kHistCopies
constexpr uint32_t kLdsByteHist = kLdsHist + kHistBins * kHistCopies; // 256 byte values, one copy
constexpr uint32_t kLdsMisc = kLdsByteHist + 256;                     // 256 words
constexpr uint32_t kLdsSeg = kLdsMisc + 256;                      // segment bytes
}  // namespace

uint32_t lds_bytes_for(uint32_t seg_bytes) { return (kLdsSeg * 4) + seg_bytes + 16; }

namespace {

// misc word layout
constexpr uint32_t kMiscItem = 0;       // next segment from the work counter
constexpr uint32_t kMiscSieveCtr = 1;   // next crossing-off work items: tiers 0-3 (word 1), tier 4 (word 2)
constexpr uint32_t kMiscSplit = 3;      // stats mode: checkpoint split bit, first and second bucket (words 3-5)
constexpr uint32_t kMiscBounds = 160;   // segment tier bounds (6 words; above the per-wave records of 16 waves)
constexpr uint32_t kMiscWave = 8;       // 8 words per wave below
// per wave: [0] first, [1] last, [2] count, [3] c41, [4] c31, [5] maxgap,
//           [6] lower count, [7] lower c41 ... (lower c31 in [kMiscWave + 8*NW])

struct SegGeom {
  uint64_t k0;
  uint32_t nbytes;
  uint32_t vlo, vhi;      // valid bit range [vlo, vhi)
  uint64_t lo_val, hi_val;
  uint32_t query;
  uint32_t from_cache;    // batch mode: bytes come from A.cache, no crossing-off
};

// Gaps between consecutive primes inside one bitmap byte x, in order, as
// 4-bit (gap/2) nibbles; they depend on x alone.
__device__ __forceinline__ uint32_t internal_gaps(uint32_t x) {
  uint32_t nib = 0, sh = 0;
  int prev = -1;
  for (int o = 0; o < 8; ++o) {
    if ((x >> o) & 1) {
      if (prev >= 0) { nib |= ((w30_const(o) - w30_const(prev)) >> 1) << sh; sh += 4; }
      prev = o;
    }
  }
  return nib;
}

// Initialise tables in LDS (once per workgroup).
__device__ void init_tables(uint32_t* lds) {
  const uint32_t tid = threadIdx.x;
  uint8_t* tab = reinterpret_cast<uint8_t*>(lds + kLdsTab);
  if (tid < 64) {
    // r = index of p mod 30 in W30, c = class (multiplier m = W30[c] mod 30)
    uint32_t r = tid >> 3, c = tid & 7;
    uint32_t prod = w30_const(r) * w30_const(c);
    uint32_t res = prod % 30, bit = 0;
    for (int o = 0; o < 8; ++o) if (w30_const(o) == res) bit = o;
    tab[tid] = (uint8_t)(prod / 30);                                     // FRAC
    tab[64 + tid] = (uint8_t)~(1u << bit);                               // NOTBIT: the class bit cleared
    tab[128 + tid] = (uint8_t)((w30_const(c) + 30 - w30_const(r)) % 30); // DM
  }
}

// The sieving-prime index bounds of a segment's tiers and plow, the least
// prime for which the p*p rule can bite (see sieve_segment).  Computed once
// per segment by one thread in init_segment and read from LDS by all.
struct SegBounds {
  uint32_t np, n0, n1, n2, n3, plow;
};
__device__ __forceinline__ SegBounds segment_bounds(const SieveArgs& A, const SegGeom& g) {
  const uint64_t k0 = g.k0;
  const uint32_t nbytes = g.nbytes;
  SegBounds B;
  uint64_t top = 30 * (k0 + nbytes) - 1;
  if (g.hi_val < top) top = g.hi_val;
  const uint64_t sq = isqrt_u64(top);
  uint64_t ti = sq / kPiTabStep + 1;
  if (ti >= A.pi_tab_len) ti = A.pi_tab_len - 1;
  const uint32_t np = B.np = A.pi_tab[ti];
  const uint32_t n0 = B.n0 = min(np, A.tier0_end);
  const uint32_t n1 = B.n1 = max(n0, min(np, A.tier1_end));
  // the sparse tier is exact for every p >= nbytes, so a short segment (a
  // small range query) moves it down: primes past the pi-table bin that
  // holds nbytes all exceed nbytes
  const uint32_t nshort = A.pi_tab[min((nbytes + kPiTabStep - 1) / kPiTabStep, A.pi_tab_len - 1)];
  const uint32_t n3 = B.n3 = max(n1, min(np, min(A.tier3_end, nshort)));
  B.n2 = max(n1, min(n3, A.tier2_end));
  // the p*p rule can bite exactly for p >= plow, the least p with
  // p * (p + 30) > 30 * k0 (one test per prime instead of a 64-bit product)
  {
    const uint64_t t30 = 30 * k0;
    uint64_t x = (uint64_t)((__builtin_sqrt(900.0 + 4.0 * (double)t30) - 30.0) * 0.5);
    while (x > 0 && (x - 1) * (x + 29) > t30) --x;
    while (x * (x + 30) <= t30) ++x;
    B.plow = (uint32_t)min(x, (uint64_t)0xFFFFFFFFu);
  }

  (void)n0;
  (void)n3;
  return B;
}

// Initialise the segment in LDS from the presieve pattern (multiples of
// 7..19 removed), zero-filling up to the next 1 KiB so later passes read whole
// 16-byte chunks.  Bytes 0..3 of the global bitmap (values below 120) get
// their true bits: 1 is not prime, and 7..113 must stay set.  Also clears the
// crossing-off work counters and publishes the segment's tier bounds.
template <int BS>
__device__ void init_segment(const SieveArgs& A, uint32_t* lds, const SegGeom& g) {
  uint8_t* seg8 = reinterpret_cast<uint8_t*>(lds + kLdsSeg);
  const uint32_t off0 = (uint32_t)(g.k0 % kPresievePeriods[0]);
  const uint32_t nchunks = ((g.nbytes + 1023) & ~1023u) >> 4;
  if (threadIdx.x < 2) lds[kLdsMisc + kMiscSieveCtr + threadIdx.x] = 0;  // sieve_segment's work counters
  // thread 0 computes the tier bounds first, so their table loads overlap its
  // pattern loads, and stores them after its chunks
  SegBounds sb;
  if (threadIdx.x == 0) sb = segment_bounds(A, g);
  // Whole 256-byte blocks go straight from the pattern to LDS
  // (global_load_lds: no registers, every block of a wave in flight at
  // once), from the copy whose byte phase makes the source word-aligned.
  // Threads fill the rest: block 0 of segment 0 (its true bits patched in),
  // the partial last block and the zero fill up to the next 1 KiB.
  const uint32_t ph = off0 & 3;
  const uint32_t* src = A.presieve + A.presieve_off[0] + ph * A.presieve_copy_words[0] + ((off0 - ph) >> 2);
  const uint32_t nfull = g.nbytes >> 8;
  const uint32_t first = (g.k0 < 4 && nfull > 0) ? 1u : 0u;  // else the thread pass covers chunk 0
  {
    const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // wave 0 may still be merging the previous segment: waves 1.. copy
    if (wave > 0)
      for (uint32_t blk = first + wave - 1; blk < nfull; blk += BS / 64 - 1)
        __builtin_amdgcn_global_load_lds((const void*)(src + 64 * blk + lane),
                                         (__attribute__((address_space(3))) void*)((lds_u32*)(lds + kLdsSeg) + 64 * blk),
                                         4, 0, 0);
  }
  const uint32_t nhead = first ? 16u : 0u;                 // 16-byte chunks of block 0
  const uint32_t nrest = nhead + (nchunks - 16 * nfull);   // plus the tail chunks
  for (uint32_t r = threadIdx.x; r < nrest; r += BS) {
    const uint32_t t = r < nhead ? r : 16 * nfull + (r - nhead);
    uint4 v;
    v.x = src[4 * t]; v.y = src[4 * t + 1]; v.z = src[4 * t + 2]; v.w = src[4 * t + 3];
    if (g.k0 < 4 && t == 0) {  // global bytes k0..3 hold the presieved primes themselves
      const uint32_t sh = 8 * (uint32_t)g.k0;
      const uint32_t m = 0xFFFFFFFFu >> sh;
      v.x = (v.x & ~m) | (A.first_word >> sh);
    }
    const uint32_t b0 = 16 * t;
    if (b0 + 16 > g.nbytes) {
      uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t wb = b0 + 4 * i;
        const uint32_t keep = (wb >= g.nbytes) ? 0u : (g.nbytes - wb >= 4 ? 4u : g.nbytes - wb);
        w[i] &= keep >= 4 ? 0xFFFFFFFFu : ((1u << (8 * keep)) - 1u);
      }
      v.x = w[0]; v.y = w[1]; v.z = w[2]; v.w = w[3];
    }
    *reinterpret_cast<uint4*>(seg8 + b0) = v;
  }
  __builtin_amdgcn_s_waitcnt(0);  // the direct LDS loads are counted as vector memory loads
  if (threadIdx.x == 0) {
    uint32_t* gb = lds + kLdsMisc + kMiscBounds;
    gb[0] = sb.np; gb[1] = sb.n0; gb[2] = sb.n1; gb[3] = sb.n2; gb[4] = sb.n3; gb[5] = sb.plow;
  }
}

// Load the segment from the resident cache bitmap (sieved bytes from k = 0)
// instead of the presieve pattern: the bytes are final, no crossing-off
// follows.  As init_segment, it zero-fills up to the next 1 KiB and clears
// bytes past nbytes; a chunk reads at most 20 bytes from its start, and only
// chunks that begin inside the segment read at all.
template <int BS>
__device__ void cache_segment(const SieveArgs& A, uint32_t* lds, const SegGeom& g) {
  uint8_t* seg8 = reinterpret_cast<uint8_t*>(lds + kLdsSeg);
  const uint32_t nchunks = ((g.nbytes + 1023) & ~1023u) >> 4;
  const uint32_t sh = (uint32_t)(g.k0 & 3);
  const uint32_t* src = reinterpret_cast<const uint32_t*>(A.cache) + (g.k0 >> 2);
  for (uint32_t t = threadIdx.x; t < nchunks; t += BS) {
    const uint32_t b0 = 16 * t;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (b0 < g.nbytes) {
      const uint32_t* p = src + 4 * t;
      const uint32_t a0 = p[0], a1 = p[1], a2 = p[2], a3 = p[3], a4 = p[4];
      v.x = __builtin_amdgcn_alignbyte(a1, a0, sh);
      v.y = __builtin_amdgcn_alignbyte(a2, a1, sh);
      v.z = __builtin_amdgcn_alignbyte(a3, a2, sh);
      v.w = __builtin_amdgcn_alignbyte(a4, a3, sh);
      if (b0 + 16 > g.nbytes) {
        uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const uint32_t wb = b0 + 4 * i;
          const uint32_t keep = (wb >= g.nbytes) ? 0u : (g.nbytes - wb >= 4 ? 4u : g.nbytes - wb);
          w[i] &= keep >= 4 ? 0xFFFFFFFFu : ((1u << (8 * keep)) - 1u);
        }
        v.x = w[0]; v.y = w[1]; v.z = w[2]; v.w = w[3];
      }
    }
    *reinterpret_cast<uint4*>(seg8 + b0) = v;
  }
}

// The read patterns (multiples of 23..109 cleared) over one segment: the 16
// bytes that AND into the segment's chunk at byte boff (a multiple of 16),
// each from the copy whose phase makes the source word-aligned.  In a
// segment starting below global byte 4, chunk 0 keeps its first 4 - k0
// bytes as they are: init_segment wrote the true primes there (23..109 among
// them).
constexpr int kReadPatterns = kPresieveGroups - 1;
struct ReadPatterns {
  const uint32_t* src[kReadPatterns];  // word of pattern 1 + i that aligns with segment byte 0
  uint32_t head_keep;                  // chunk 0, word 0: bits left untouched
};
__device__ __forceinline__ ReadPatterns read_patterns(const SieveArgs& A, uint64_t k0) {
  ReadPatterns P;
#pragma unroll
  for (int i = 0; i < kReadPatterns; ++i) {
    const uint32_t off = (uint32_t)(k0 % kPresievePeriods[1 + i]);
    const uint32_t ph = off & 3;
    P.src[i] = A.presieve + A.presieve_off[1 + i] + ph * A.presieve_copy_words[1 + i] + ((off - ph) >> 2);
  }
  P.head_keep = k0 < 4 ? (0xFFFFFFFFu >> (8 * (uint32_t)k0)) : 0u;
  return P;
}
__device__ __forceinline__ uint4 read_pattern_chunk(const ReadPatterns& P, uint32_t boff) {
  uint4 v = make_uint4(~0u, ~0u, ~0u, ~0u);
#pragma unroll
  for (int i = 0; i < kReadPatterns; ++i) {
    const uint32_t* p = P.src[i] + (boff >> 2);
    v.x &= p[0]; v.y &= p[1]; v.z &= p[2]; v.w &= p[3];
  }
  if (boff == 0) v.x |= P.head_keep;
  return v;
}
__device__ __forceinline__ void and4(uint4& q, const uint4& b) {
  q.x &= b.x; q.y &= b.y; q.z &= b.z; q.w &= b.w;
}
// The read patterns' words for the chunk at boff, loaded but not yet
// combined, so a load issued one step ahead is waited for only where the
// next step uses it.
struct PatternWords {
  uint4 w[kReadPatterns];
};
__device__ __forceinline__ PatternWords load_pattern_words(const ReadPatterns& P, uint32_t boff) {
  PatternWords W;
#pragma unroll
  for (int i = 0; i < kReadPatterns; ++i) {
    const uint32_t* p = P.src[i] + (boff >> 2);
    W.w[i] = make_uint4(p[0], p[1], p[2], p[3]);
  }
  return W;
}
__device__ __forceinline__ void and_pattern_words(uint4& q, const PatternWords& W, const ReadPatterns& P,
                                                  uint32_t boff) {
  uint4 v = W.w[0];
#pragma unroll
  for (int i = 1; i < kReadPatterns; ++i) and4(v, W.w[i]);
  if (boff == 0) v.x |= P.head_keep;
  and4(q, v);
}

// A load that every lane makes from the same address of memory no kernel
// writes: through the constant address space it is a scalar load (scalar
// cache, waited for apart from the vector loads in flight).
__device__ __forceinline__ uint32_t load_uniform(const uint32_t* p, uint32_t i) {
  return ((const __attribute__((address_space(4))) uint32_t*)p)[i];
}

__device__ __forceinline__ void lds_and_at(uint32_t addr, uint32_t m) {
  __hip_atomic_fetch_and((lds_u32*)(uintptr_t)addr, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS byte address of the segment bitmap.
__device__ __forceinline__ uint32_t seg_base_addr(uint32_t*) {
  // dynamic LDS starts at address 0 (prepare_kernels checks that no static
  // LDS precedes it), so the base is a constant the ds instructions take in
  // their offset field
  return kLdsSeg * 4;
}

// Crossing masks.  mc = 0xFFFFFF00 | nb, nb the byte ~(1 << bit) of a class
// (table row NOTBIT).  The mask that clears the class bit of the byte at
// segment offset off is mc rotated left by 8 (off mod 4) bits, i.e. right by
// 8 (-off mod 4): one alignbyte on -off (a shift off * (2^24 - 8) for
// alignbit would be lowered to a quarter-rate 32-bit multiply, off not being
// known to fit 24 bits).  A prime's next crossing is p bytes on, so its mask
// is the last one rotated by the same amount for p: alignbit rotates right
// by its shift's low five bits, and p * (2^24 - 8) (a 24-bit multiply,
// p < 2^24) has low five bits -8 p mod 32.
__device__ __forceinline__ uint32_t rot_step(uint32_t x) { return __umul24(x, 0xFFFFF8u); }
__device__ __forceinline__ uint32_t rotate_by(uint32_t m, uint32_t step) {
  return __builtin_amdgcn_alignbit(m, m, step);
}
__device__ __forceinline__ uint32_t byte_mask(uint32_t mc, uint32_t off) {
  return __builtin_amdgcn_alignbyte(mc, mc, 0u - off);
}
// mc of class c from a prime's 8-byte NOTBIT row (v_perm: bytes 1-3 0xFF)
__device__ __forceinline__ uint32_t row_mask(uint64_t row, int c) {
  return __builtin_amdgcn_perm((uint32_t)(row >> 32), (uint32_t)row, 0x0D0D0D00u | (uint32_t)c);
}
__device__ __forceinline__ uint32_t mask_word(uint32_t nb) { return 0xFFFFFF00u | nb; }

// Clear the class bit (mask word mc) of bytes j, j+p, j+2p, ... below `end`
// (segment byte offsets; base = LDS byte address of the segment).  The
// byte-in-dword pattern repeats every 4 crossings, so the four dword deltas
// and masks are loop invariants: one add per crossing plus the ds_and_b32.
// FULL: the caller knows a first group fits (j + 3p < end).
template <bool FULL = false>
__device__ __forceinline__ void cross_class(uint32_t base, uint32_t j, uint32_t p, uint32_t end, uint32_t mc) {
  const uint32_t j0 = j & 3;
  const uint32_t d1 = (j0 + p) & ~3u, d2 = (j0 + 2 * p) & ~3u, d3 = (j0 + 3 * p) & ~3u;
  const uint32_t rp = rot_step(p);
  const uint32_t m0 = byte_mask(mc, j0);
  const uint32_t m1 = rotate_by(m0, rp), m2 = rotate_by(m1, rp), m3 = rotate_by(m2, rp);
  uint32_t a = base + (j & ~3u);
  const int32_t lim = (int32_t)(base + end) - (int32_t)(j0 + 3 * p);  // a full group fits while a < lim
  const uint32_t step = 4 * p;
  if (FULL || (int32_t)a < lim) {
    do {
      lds_and_at(a, m0);
      lds_and_at(a + d1, m1);
      lds_and_at(a + d2, m2);
      lds_and_at(a + d3, m3);
      a += step;
    } while ((int32_t)a < lim);
  }
  const uint32_t jt = a - base + j0;
  if (jt < end) lds_and_at(a, m0);
  if (jt + p < end) lds_and_at(a + d1, m1);
  if (jt + 2 * p < end) lds_and_at(a + d2, m2);
}

// Clear the class bit (mask word mc) of the crossings s + (128 r + t) p,
// t = 0..3, r = 0, 1, ..., the first n of them: groups of four consecutive
// crossings, 128 crossings apart (a half-wave's 32 lanes take the 32 groups
// between).  The byte-in-dword pattern of a group is the same in every
// round, so its deltas and masks are loop invariants.
__device__ __forceinline__ void cross_rounds(uint32_t base, uint32_t s, uint32_t p, int32_t n, uint32_t mc) {
  const uint32_t j0 = s & 3;
  const uint32_t d1 = (j0 + p) & ~3u, d2 = (j0 + 2 * p) & ~3u, d3 = (j0 + 3 * p) & ~3u;
  const uint32_t rp = rot_step(p);
  const uint32_t m0 = byte_mask(mc, j0);
  const uint32_t m1 = rotate_by(m0, rp), m2 = rotate_by(m1, rp), m3 = rotate_by(m2, rp);
  uint32_t a = base + (s & ~3u);
  const uint32_t step = 128 * p;
  while (n >= 4) {
    lds_and_at(a, m0);
    lds_and_at(a + d1, m1);
    lds_and_at(a + d2, m2);
    lds_and_at(a + d3, m3);
    a += step;
    n -= 128;
  }
  if (n > 0) lds_and_at(a, m0);
  if (n > 1) lds_and_at(a + d1, m1);
  if (n > 2) lds_and_at(a + d2, m2);
}

// Clear the class bit (mask word mc) of bytes s, s + PH p, s + 2 PH p, ...
// below end: every PH-th crossing of a class (PH a multiple of 4), so the
// byte within the dword, and with it the mask, never changes.
template <uint32_t PH>
__device__ __forceinline__ void cross_phase(uint32_t base, uint32_t s, uint32_t p, uint32_t end, uint32_t mc) {
  static_assert(PH % 4 == 0, "the stride must keep the byte within the dword");
  const uint32_t m = byte_mask(mc, s);
  uint32_t a = base + (s & ~3u);
  const uint32_t step = PH * p;
  const int32_t lim = (int32_t)(base + end - (s & 3u));  // a crossing at dword a exists while a < lim
  while ((int32_t)(a + 3 * step) < lim) {
    lds_and_at(a, m);
    lds_and_at(a + step, m);
    lds_and_at(a + 2 * step, m);
    lds_and_at(a + 3 * step, m);
    a += 4 * step;
  }
  if ((int32_t)a < lim) lds_and_at(a, m);
  if ((int32_t)(a + step) < lim) lds_and_at(a + step, m);
  if ((int32_t)(a + 2 * step) < lim) lds_and_at(a + 2 * step, m);
}

// floor(x / p) for x < 2^22, 0 < p < 2^24 (float reciprocal, one correction each way).
__device__ __forceinline__ uint32_t div_u32(uint32_t x, uint32_t p) {
  uint32_t q = (uint32_t)((float)x * __builtin_amdgcn_rcpf((float)p));
  if (q * p > x) --q;
  if ((q + 1) * p <= x) ++q;
  return q;
}

// First crossing of class c at or after byte k0 of the global bitmap,
// relative to k0; never below p*p (the reference rule, context.cu:354-388).
__device__ __forceinline__ uint64_t first_crossing(uint32_t p, uint32_t r, uint32_t c, uint32_t rk, uint64_t k0,
                                                   const lds_u8* tab) {
  const uint32_t cc = (p / 30) * w30(c) + tab[r * 8 + c];
  int32_t off = (int32_t)cc - (int32_t)rk;
  if (off < 0) off += (int32_t)p;
  uint64_t offl = (uint32_t)off;
  if (30 * k0 < (uint64_t)p * p + 30ull * p) {
    const uint64_t m = (uint64_t)p + tab[128 + r * 8 + c];
    const uint64_t fb = (uint64_t)p * m / 30;
    if (k0 + offl < fb) offl = fb - k0;
  }
  return offl;
}

// Clear the class bit (mask word mc) of segment byte off.
__device__ __forceinline__ void cross_one(uint32_t base, uint32_t off, uint32_t mc) {
  lds_and_at(base + (off & ~3u), byte_mask(mc, off));
}

// The crossings of a prime p >= nbytes (at most one per class), every class
// checked in turn with no search and no loop: off = (cc_c - rk) mod p.
__device__ __forceinline__ void cross_classes_once(uint32_t base, uint32_t p, uint32_t r, uint32_t rk,
                                                   uint32_t nbytes, const lds_u8* tab) {
  const uint32_t q30 = p / 30;
  const uint64_t frac = *(const lds_u64*)(tab + r * 8);
  const uint64_t bits = *(const lds_u64*)(tab + 64 + r * 8);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    // cm = cc - rk (mod 2^32) as a 24-bit multiply-add on the byte's
    // difference from rk, then off = (cc - rk) mod p = min(cm, cm + p)
    const uint32_t cm = __umul24(q30, w30_const(c)) + ((uint32_t)((frac >> (8 * c)) & 0xFF) - rk);
    const uint32_t off = min(cm, cm + p);
    if (off < nbytes) cross_one(base, off, row_mask(bits, c));
  }
}

// Every crossing of p (rk = k0 mod p) in the segment, class by class, from
// two packed 64-bit LDS rows: off = min(cc + p - rk, cc - rk), the p*p rule
// applied if LOW.  LONG walks each class with the unrolled cross_class loop
// (many crossings per class), otherwise one at a time.  Callers pass LOW for
// a whole wave when the rule can bite for any of its lanes (p >= plow): the
// check leaves off unchanged where p < plow, whose p * (p + 30) <= 30 k0, so
// the wave takes one scalar branch instead of an exec-masked one per class.
template <bool LONG, bool LOW, bool HITS = false>
__device__ __forceinline__ void cross_prime_classes(uint32_t base, uint32_t p, uint32_t r, uint32_t rk, uint64_t k0,
                                                    uint32_t nbytes, const lds_u8* tab, uint32_t plow) {
  const uint32_t q30 = p / 30;
  const uint64_t frac = *(const lds_u64*)(tab + r * 8);
  const uint64_t bits = *(const lds_u64*)(tab + 64 + r * 8);
  const uint32_t rp = rot_step(p);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t cm = __umul24(q30, w30_const(c)) + ((uint32_t)((frac >> (8 * c)) & 0xFF) - rk);
    uint32_t off = min(cm, cm + p);  // (cc - rk) mod p, as unsigned wrap-around
    if (LOW) {
      const uint64_t m = (uint64_t)p + tab[128 + r * 8 + c];
      const uint64_t fb = (uint64_t)p * m / 30;
      if (k0 + off < fb) off = (fb - k0 > 0xFFFFFFFFull) ? 0xFFFFFFFFu : (uint32_t)(fb - k0);
    }
    const uint32_t mc = row_mask(bits, c);
    if (LONG) {
      // (HITS: every lane's 4p <= nbytes and no p*p rule, so off + 3p < nbytes)
      if (HITS || off < nbytes) cross_class<HITS>(base, off, p, nbytes, mc);
    } else if (HITS || off < nbytes) {
      // one crossing per step: off advances by p, so the byte within the
      // dword advances by p mod 4 and the mask rotates by 8 (p mod 4) bits
      // (HITS: every lane's p < nbytes and no p*p rule, so off < p < nbytes)
      uint32_t m = byte_mask(mc, off);
      do {
        lds_and_at(base + (off & ~3u), m);
        off += p;
        m = rotate_by(m, rp);
      } while (off < nbytes);
    }
  }
}

// The crossings of a prime with MINK <= K <= MINK + 1 crossings in every
// class (MINK p <= nbytes <= (MINK + 1) p, no p*p rule): MINK unconditional,
// the last one predicated, no loop.
template <int MINK>
__device__ __forceinline__ void cross_prime_banded(uint32_t base, uint32_t p, uint32_t r, uint32_t rk,
                                                   uint32_t nbytes, const lds_u8* tab) {
  const uint32_t q30 = p / 30;
  const uint64_t frac = *(const lds_u64*)(tab + r * 8);
  const uint64_t bits = *(const lds_u64*)(tab + 64 + r * 8);
  const uint32_t rp = rot_step(p);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint32_t cm = __umul24(q30, w30_const(c)) + ((uint32_t)((frac >> (8 * c)) & 0xFF) - rk);
    uint32_t off = min(cm, cm + p);
    uint32_t m = byte_mask(row_mask(bits, c), off);
#pragma unroll
    for (int k = 0; k < MINK; ++k) {
      lds_and_at(base + (off & ~3u), m);
      off += p;
      m = rotate_by(m, rp);
    }
    if (off < nbytes) lds_and_at(base + (off & ~3u), m);
  }
}

// Cross off every sieving prime >= 23 with p*p <= the segment's top value.
//   tier 0 (p < 160): a half-wave per (prime, class), its 32 lanes taking
//     the class's crossings in rounds whose LDS banks never collide;
//   tier 1 (p below the tier1_end-th prime): two primes per wave turn, one
//     per half-wave, whose 32 lanes are 8 classes x 4 phases, lane (c, t)
//     taking every fourth crossing of class c from the t-th, so its mask
//     never changes;
//   tier 2 (p < seg_bytes/5): one prime per lane, each class walked by the
//     unrolled loop;
//   tier 3 (seg_bytes/5 <= p < seg_bytes): one prime per lane, at most five
//     crossings per class: where every lane's prime has K or K + 1 in each
//     class (a band between nbytes/(K+1) and nbytes/K, which a chunk of
//     consecutive primes nearly always lies in) K unconditional crossings and
//     one predicated (cross_prime_banded), else a loop per class
//     (cross_prime_classes);
//   tier 4 (p >= seg_bytes): one prime per lane, all eight classes checked
//     for their single possible crossing (cross_classes_once).
template <int BS>
__device__ void sieve_segment(const SieveArgs& A, uint32_t* lds, const SegGeom& g) {
  constexpr int NW = BS / 64;
  const uint32_t lane = threadIdx.x & 63;
  // the tables' LDS address is a constant (dynamic LDS starts at 0), so it
  // folds into the instructions' offset fields
  const lds_u8* tab = (const lds_u8*)(uintptr_t)(kLdsTab * 4);
  const uint32_t base = seg_base_addr(lds);
  const uint64_t k0 = g.k0;
  const uint32_t nbytes = g.nbytes;

  const uint32_t* gb = lds + kLdsMisc + kMiscBounds;
  const uint32_t np = gb[0], n0 = gb[1], n1 = gb[2], n2 = gb[3], n3 = gb[4], plow = gb[5];

  // Work items of tiers 0-3, heaviest first: the 64-prime chunks of tier 2
  // (its first chunk alone is a large share of a wave's work), tier-0
  // primes, tier-1 prime pairs, then the 64-prime chunks of tier 3.  Each wave takes its next item from an LDS counter (cleared by
  // init_segment), so waves that drew heavy items take fewer and all of them
  // reach tier 4 together.  Tier 4 has its own counter below.
  const uint32_t e2 = (n2 - n1 + 63) >> 6;
  const uint32_t e0 = e2 + n0;
  const uint32_t e1 = e0 + (n1 - n0 + 1) / 2;
  const uint32_t e3 = e1 + ((n3 - n2 + 63) >> 6);
  auto take = [&]() -> uint32_t {
    uint32_t v = 0;
    if (lane == 0)
      v = __hip_atomic_fetch_add((lds_u32*)(uintptr_t)((kLdsMisc + kMiscSieveCtr) * 4), 1u, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
    return __builtin_amdgcn_readlane(v, 0);
  };
  // this lane's table entry for a chunk item (0: none); loaded one item
  // ahead so the read's latency overlaps the current item's crossings
  auto lane_entry = [&](uint32_t it) -> uint32_t {
    uint32_t i = 0xFFFFFFFFu, end = 0;
    if (it < e2) { i = n1 + 64 * it + lane; end = n2; }
    else if (it >= e1 && it < e3) { i = n2 + 64 * (it - e1) + lane; end = n3; }
    return i < end ? A.primes[i] : 0u;
  };
  // Taking order: one tier-2 chunk per wave first, then the other tier-2
  // chunks spread evenly among the tier-0/1 items (position j of that run is
  // a chunk when floor((j+1) Am / Mm) > floor(j Am / Mm)), then tier 3, so
  // waves on the setup-bound one-prime items overlap waves on LDS-bound
  // chunks.  remap turns the counter's value into the item index above.
  const uint32_t H = min(e2, (uint32_t)NW);
  const uint32_t Am = e2 - H, Mm = e1 - H;
  // (the float quotient below is exact to within one for products below
  // 2^24, far above any tier layout the engine configures)
  const bool spread = Am != 0 && (uint64_t)Am * Mm < (1u << 24);
  auto remap = [&](uint32_t t) -> uint32_t {
    if (t < H || t >= e1 || !spread) return t;
    const uint32_t j = t - H;
    auto fl = [&](uint32_t x) -> uint32_t {  // floor(x * Am / Mm), exact
      const uint32_t y = x * Am;
      uint32_t q = (uint32_t)((float)y * __builtin_amdgcn_rcpf((float)Mm));
      if (q * Mm > y) --q;
      if ((q + 1) * Mm <= y) ++q;
      return q;
    };
    const uint32_t a0 = fl(j), a1 = fl(j + 1);
    return a1 > a0 ? H + a0 : e2 + (j - a0);
  };
  uint32_t it = remap(take());
  uint32_t e = lane_entry(it);
  // the first entry is waited for here, so that inside the loop only the
  // next item's entry is in flight and its wait falls at the loop's end
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  while (it < e3) {
    const uint32_t nit = remap(take());
    const uint32_t ne = lane_entry(nit);
    const uint32_t p = e & 0xFFFFFFu, r = e >> 24;
    if (it < e2) {
      // tier 2: one prime per lane, each class walked by the unrolled loop
      if (e) {
        const uint32_t rk = mod_u64(k0, p);
        if (__builtin_amdgcn_ballot_w64(p >= plow)) cross_prime_classes<true, true>(base, p, r, rk, k0, nbytes, tab, plow);
        else if (__builtin_amdgcn_ballot_w64(4 * p > nbytes)) cross_prime_classes<true, false>(base, p, r, rk, k0, nbytes, tab, plow);
        else cross_prime_classes<true, false, true>(base, p, r, rk, k0, nbytes, tab, plow);
      }
    } else if (it < e0) {
      // tier 0: item = prime; half-wave h takes classes h, h + 2, h + 4, h + 6
      // in turn (k0 mod p computed once).  A class's K crossings go in
      // rounds of 128: lane l takes crossings
      // 128 r + 4 l .. 128 r + 4 l + 3, so in each instruction lane l is l*p
      // dwords after lane 0: p being odd, the 32 lanes of a half-wave always
      // address 32 different LDS banks.
      const uint32_t t0 = it - e2;
      const uint32_t hl = lane & 31, half = lane >> 5;
      const uint32_t e0i = load_uniform(A.primes, t0);
      const uint32_t p0 = e0i & 0xFFFFFFu, r0 = e0i >> 24;
      const uint32_t rk0 = mod_u64(k0, p0);
      for (uint32_t q = 0; q < 4; ++q) {
        const uint32_t c = (q << 1) | half;
        const uint64_t offl = first_crossing(p0, r0, c, rk0, k0, tab);
        if (offl < nbytes) {
          const uint32_t j = (uint32_t)offl;
          const uint32_t K = div_u32(nbytes - 1 - j, p0) + 1;  // crossings j + k*p < nbytes
          if (4 * hl < K) cross_rounds(base, j + 4 * hl * p0, p0, (int32_t)(K - 4 * hl), mask_word(tab[64 + r0 * 8 + c]));
        }
      }
    } else if (it < e1) {
      // tier 1: two primes per wave, one per half-wave, whose lanes are 8
      // classes x 4 phases: lane (c, t) takes crossings t, t + 4, t + 8, ...
      // of class c
      const uint32_t c = lane & 7, sr = (lane >> 3) & 3, half = lane >> 5;
      const uint32_t i1 = n0 + 2 * (it - e0);
      const uint32_t ea = load_uniform(A.primes, i1);
      const uint32_t eb = i1 + 1 < n1 ? load_uniform(A.primes, i1 + 1) : 0u;
      const uint32_t e1i = half ? eb : ea;
      if (e1i) {
        const uint32_t p1 = e1i & 0xFFFFFFu, r1 = e1i >> 24;
        const uint64_t offl = first_crossing(p1, r1, c, mod_u64(k0, p1), k0, tab) + (uint64_t)sr * p1;
        if (offl < nbytes) cross_phase<4>(base, (uint32_t)offl, p1, nbytes, mask_word(tab[64 + r1 * 8 + c]));
      }
    } else {
      // tier 3: one prime per lane, at most five crossings per class; the
      // wave takes a band's loop-free form when all its primes lie in it
      if (e) {
        const uint32_t rk = mod_u64(k0, p);
        if (__builtin_amdgcn_ballot_w64(p >= plow)) cross_prime_classes<false, true>(base, p, r, rk, k0, nbytes, tab, plow);
        else if (__builtin_amdgcn_ballot_w64(p >= nbytes)) cross_prime_classes<false, false>(base, p, r, rk, k0, nbytes, tab, plow);
        else if (!__builtin_amdgcn_ballot_w64(4 * p > nbytes || 5 * p < nbytes)) cross_prime_banded<4>(base, p, r, rk, nbytes, tab);
        else if (!__builtin_amdgcn_ballot_w64(3 * p > nbytes || 4 * p < nbytes)) cross_prime_banded<3>(base, p, r, rk, nbytes, tab);
        else if (!__builtin_amdgcn_ballot_w64(2 * p > nbytes || 3 * p < nbytes)) cross_prime_banded<2>(base, p, r, rk, nbytes, tab);
        else if (!__builtin_amdgcn_ballot_w64(2 * p < nbytes)) cross_prime_banded<1>(base, p, r, rk, nbytes, tab);
        else cross_prime_classes<false, false, true>(base, p, r, rk, k0, nbytes, tab, plow);
      }
    }
    it = nit;
    e = ne;
  }
  // tier 4 (p >= seg_bytes): at most one crossing per class, each of the
  // eight classes checked in turn, fully unrolled and predicated (a search
  // for the run of classes that hit, with a loop over it, cost more: its
  // divergent loops and their exec-mask upkeep outweighed the classes it
  // skipped, even for primes 20 segments long).  Items are groups of g4
  // 64-prime chunks (about eight per wave: fine enough to balance the waves,
  // few enough to bound the counter traffic of a short segment with
  // thousands of chunks)
  // taken from a second LDS counter; the primes of the next chunk, and of
  // the next item's first chunk, are loaded ahead.
  const uint32_t c4 = (np - n3 + 63) >> 6;
  const uint32_t g4 = max(1u, (c4 + 8 * NW - 1) / (8 * NW));
  const uint32_t items4 = (c4 + g4 - 1) / g4;
  auto take4 = [&]() -> uint32_t {
    uint32_t v = 0;
    if (lane == 0)
      v = __hip_atomic_fetch_add((lds_u32*)(uintptr_t)((kLdsMisc + kMiscSieveCtr + 1) * 4), 1u, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
    return __builtin_amdgcn_readlane(v, 0);
  };
  auto first4 = [&](uint32_t t) -> uint32_t {
    const uint32_t i = n3 + 64 * g4 * t + lane;
    return (t < items4 && i < np) ? A.primes[i] : 0u;
  };
  uint32_t it4 = take4();
  uint32_t f4 = first4(it4);
  while (it4 < items4) {
    const uint32_t nit4 = take4();
    const uint32_t nf4 = first4(nit4);
    const uint32_t b0 = n3 + 64 * g4 * it4;
    const uint32_t b1 = min(np, b0 + 64 * g4);
    uint32_t e_next = f4;
    for (uint32_t b = b0; b < b1; b += 64) {
      const uint32_t i = b + lane;
      const uint32_t e4 = e_next;
      e_next = (i + 64 < b1) ? A.primes[i + 64] : 0u;
      if (i >= b1) continue;
      const uint32_t p = e4 & 0xFFFFFFu, r = e4 >> 24;
      const uint32_t rk = mod_u64(k0, p);
      if (p >= plow) {  // the p*p rule can bite: general path
        cross_prime_classes<false, true>(base, p, r, rk, k0, nbytes, tab, plow);
        continue;
      }
      cross_classes_once(base, p, r, rk, nbytes, tab);
    }
    it4 = nit4;
    f4 = nf4;
  }
}

__device__ __forceinline__ void hist_add_lds(uint32_t* lds, uint32_t h, uint32_t copy) {
  lds_add(lds + kLdsHist + h * kHistCopies + copy, 1u);
}

// Mask a 16-byte chunk (bit offset cb) to the valid bit range [vlo, vhi).
__device__ __forceinline__ void mask_chunk(uint4& q, uint32_t cb, uint32_t vlo, uint32_t vhi) {
  uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t b = cb + 32 * i;
    uint32_t m = 0xFFFFFFFFu;
    if (vlo > b) m = (vlo - b >= 32) ? 0u : (m << (vlo - b));
    if (vhi < b + 32) {
      uint32_t keep = (vhi <= b) ? 0u : vhi - b;
      m &= keep == 0 ? 0u : (0xFFFFFFFFu >> (32 - keep));
    }
    w[i] &= m;
  }
  q.x = w[0]; q.y = w[1]; q.z = w[2]; q.w = w[3];
}

// Bits of a 16-byte chunk (bit offset cb) below bit index split.
__device__ __forceinline__ uint4 below_split(uint4 q, uint32_t cb, uint32_t split) {
  uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t b = cb + 32 * i;
    uint32_t keep = split <= b ? 0u : (split - b >= 32 ? 32u : split - b);
    w[i] &= keep == 0 ? 0u : (0xFFFFFFFFu >> (32 - keep));
  }
  return make_uint4(w[0], w[1], w[2], w[3]);
}

__device__ __forceinline__ uint32_t popc4(uint4 q) {
  return __builtin_popcount(q.x) + __builtin_popcount(q.y) + __builtin_popcount(q.z) + __builtin_popcount(q.w);
}
__device__ __forceinline__ uint32_t popc4m(uint4 q, uint32_t m) {
  return __builtin_popcount(q.x & m) + __builtin_popcount(q.y & m) + __builtin_popcount(q.z & m) +
         __builtin_popcount(q.w & m);
}

// Segment-relative value of the highest set bit of a chunk at byte offset boff (0 if empty).
__device__ __forceinline__ uint32_t chunk_last(uint4 q, uint32_t boff) {
  uint32_t w = q.w, wi = 3;
  if (!w) { w = q.z; wi = 2; }
  if (!w) { w = q.y; wi = 1; }
  if (!w) { w = q.x; wi = 0; }
  if (!w) return 0;
  uint32_t bit = 31 - __builtin_clz(w);
  return 30 * (boff + 4 * wi + (bit >> 3)) + w30(bit & 7);
}

// Segment-relative value of the lowest set bit of a chunk at byte offset boff (0 if empty).
__device__ __forceinline__ uint32_t chunk_first(uint4 q, uint32_t boff) {
  uint32_t w = q.x, wi = 0;
  if (!w) { w = q.y; wi = 1; }
  if (!w) { w = q.z; wi = 2; }
  if (!w) { w = q.w; wi = 3; }
  if (!w) return 0;
  const uint32_t bit = __builtin_ctz(w);
  return 30 * (boff + 4 * wi + (bit >> 3)) + w30(bit & 7);
}

// Exclusive "last prime before this lane" within one 64-lane step.
__device__ __forceinline__ uint32_t prev_in_step(uint32_t lastv, uint32_t lane, uint32_t& step_last) {
  uint32_t incl = lastv;
  const bool dense = __ballot(lastv == 0) == 0ull;
  if (dense) {
    // every lane holds a prime: the predecessor is simply lane-1's last
    // (DPP wave_shr:1: lane 0 keeps the 0 passed as the old value)
    step_last = __builtin_amdgcn_readlane(lastv, 63);
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lastv, 0x138, 0xF, 0xF, false);
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = (uint32_t)__shfl_up((int)incl, d, 64);
    if (lane >= (uint32_t)d) incl = max(incl, t);
  }
  step_last = __shfl((int)incl, 63, 64);
  uint32_t up = (uint32_t)__shfl_up((int)incl, 1, 64);
  return lane ? up : 0u;
}

template <int MODE>
__device__ __forceinline__ void note_gap(const SieveArgs& A, uint32_t* lds, uint32_t gap, uint32_t vrel,
                                          const SegGeom& g, uint32_t copy, uint32_t& maxg) {
  uint32_t h = gap >> 1;
  if (h < kHistBins) {
    hist_add_lds(lds, h, copy);
  } else if (MODE == kModeStats) {
    atomicAdd(A.hist + (h < kGapBins ? h : kGapBins - 1), 1ull);
  } else {
    atomicAdd(A.qhist + (size_t)g.query * kQueryBins, 1u);
  }
  maxg = max(maxg, gap);
  if (MODE == kModeStats && gap >= A.large_threshold) {
    unsigned int slot = atomicAdd(A.n_large, 1u);
    if (slot < A.large_cap) {
      A.large[slot].position = 30 * g.k0 + vrel;
      A.large[slot].gap = gap;
      A.large[slot].seg = 0;
    }
  }
}

// Exact handling of the rare big gaps of a lane's chunk: gaps >= 1024 were
// counted in the last LDS bin by the fast path and move to their own bin;
// gaps >= the large-gap threshold are appended to the large-gap list.
template <int MODE>
__device__ void chunk_big_gaps(const SieveArgs& A, uint32_t* lds, uint4 q, uint32_t boff, uint32_t prev,
                               const SegGeom& g, uint32_t copy) {
  const uint32_t ws[4] = {q.x, q.y, q.z, q.w};
  for (int wi = 0; wi < 4; ++wi) {
    uint32_t w = ws[wi];
    while (w) {
      const uint32_t b = __builtin_ctz(w);
      w &= w - 1;
      const uint32_t v = 30 * (boff + 4 * wi + (b >> 3)) + w30(b & 7);
      if (prev) {
        const uint32_t gap = v - prev;
        if (MODE == kModeStats && (gap >> 1) >= kHistBins) {
          lds_add(lds + kLdsHist + (kHistBins - 1) * kHistCopies + copy, 0xFFFFFFFFu);  // undo the clamp
          const uint32_t h = gap >> 1;
          atomicAdd(A.hist + (h < kGapBins ? h : kGapBins - 1), 1ull);
        }
        if (MODE == kModeStats && gap >= A.large_threshold) {
          const unsigned int slot = atomicAdd(A.n_large, 1u);
          if (slot < A.large_cap) {
            A.large[slot].position = 30 * g.k0 + v;
            A.large[slot].gap = gap;
            A.large[slot].seg = 0;
          }
        }
      }
      prev = v;
    }
  }
}

// In-order walk of the sieved segment: counts, tallies, gaps; per-wave
// summaries into LDS misc (thread 0 merges them afterwards).
//
// Each lane takes 16 bytes per step (one ds_read_b128; a wave covers 1 KiB).
// Per byte, branch-free: the gap from the previous prime to the byte's first
// prime goes to the LDS gap histogram; the gaps between primes of the same
// byte depend only on the byte value, so a byte with >= 2 primes bumps a
// histogram of byte values instead, converted to gaps when flushed.
//
// Positions inside the byte loop are 8 * (segment-relative value): a gap g
// then is 8g, half the byte offset 32 * (g / 2) of its histogram bin (eight
// u32 copies per bin).  The wheel offsets of a byte's first and last prime
// come from v_perm byte tables (no LDS read).  A lane with no earlier prime in its wave starts from
// its chunk's first byte; the one spurious gap that gives (chunk start to the
// first prime) is taken back after the loop.  The chunk's running maximum
// `lm` takes the gaps that close in the chunk, each raised by a constant D
// above every zero byte's distance from the last prime; a chunk without a
// prime keeps the largest such distance instead, which never exceeds the
// gap that closes over it, so `lm` stays a bound.
template <int MODE, int BS>
__device__ void walk_segment(const SieveArgs& A, uint32_t* lds, const SegGeom& g, uint32_t split,
                             const PatternWords& bfirst) {
  constexpr int NW = BS / 64;
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint8_t* seg8 = reinterpret_cast<const uint8_t*>(lds + kLdsSeg);
  const uint32_t copy = lane & 7;
  const uint32_t wave_bytes = A.seg_bytes / NW;
  const uint32_t M41 = (g.k0 & 1) ? 0x99669966u : 0x66996699u;
  const uint32_t M31 = 0x2B2B2B2Bu;
  const bool partial = g.vlo > 0 || g.vhi < 8 * g.nbytes;
  // gaps at or above this take the exact slow path (stats mode only)
  const uint32_t slow = (MODE == kModeStats) ? min(A.large_threshold, 2 * kHistBins) : 0xFFFFFFFFu;
  // The kernels declare no static LDS (prepare_kernels checks), so the
  // dynamic array starts at LDS address 0: 4x addresses byte value x in the
  // byte-value histogram through the instruction offset field.
  lds_u8* const H8 = (lds_u8*)(lds_u32*)lds + (kLdsHist + copy) * 4;
  constexpr uint32_t kCap8 = (kHistBins - 1) * 16;  // last bin collects longer gaps
  static_assert(kHistCopies == 8, "bin stride 32 bytes = 2 * (8 * gap)");
  // D lifts every closing gap (times 8) above every zero byte's distance;
  // H8m is H8's address less 2 D, so H8m + 2 (D + 8 gap) addresses the bin
  constexpr uint32_t D = 0x08080801u;
  const uint32_t H8m = (uint32_t)(uintptr_t)H8 - 2u * D;

  uint32_t carry = 0, wfirst = 0;
  uint32_t cnt = 0, c41 = 0, c31 = 0, maxg = 0, lcnt = 0, l41 = 0, l31 = 0;
  const uint32_t wbeg = wave * wave_bytes;
  const uint32_t wend = min(wbeg + wave_bytes, g.nbytes);
  // the read patterns are ANDed in as each chunk is read (not for batch
  // segments filled from the resident cache, whose bytes are final); the
  // next step's are loaded one step ahead, the first step's (bfirst) by the
  // caller before the barrier that precedes the walk
  const bool apply_rp = !(MODE == kModeBatch && g.from_cache);  // uniform per workgroup
  const ReadPatterns RP = read_patterns(A, g.k0);
  PatternWords bnext = bfirst;
  // waves further along drop in issue priority, so lagging waves catch up
  // and the walk's waves reach the barrier together
  __builtin_amdgcn_s_setprio(3);
  for (uint32_t base = wbeg; base < wend; base += 1024) {
    {
      const uint32_t done = (base - wbeg) >> 10;
      if (done == 3) __builtin_amdgcn_s_setprio(2);
      else if (done == 6) __builtin_amdgcn_s_setprio(1);
      else if (done == 8) __builtin_amdgcn_s_setprio(0);
    }
    const uint32_t boff = base + 16 * lane;
    // (dynamic LDS starts at address 0: the segment's address is a constant)
    const u32x4 qv = *(const lds_u32x4*)(uintptr_t)(kLdsSeg * 4 + boff);
    uint4 q = make_uint4(qv.x, qv.y, qv.z, qv.w);
    if (apply_rp) {
      // combine this step's words (loaded one step ahead), then load the
      // next step's into the same registers (the last step reloads its own)
      and_pattern_words(q, bnext, RP, boff);
      bnext = load_pattern_words(RP, base + 1024 < wend ? boff + 1024 : boff);
    }
    if (partial) {
      const uint32_t cb = 8 * boff;
      if (cb < g.vlo || cb + 128 > g.vhi) mask_chunk(q, cb, g.vlo, g.vhi);
    }
    cnt += popc4(q);
    if (MODE == kModeStats) {
      c41 += popc4m(q, M41);
      c31 += popc4m(q, M31);
      if (split != 0xFFFFFFFFu) {
        const uint4 lo = below_split(q, 8 * boff, split);
        lcnt += popc4(lo);
        l41 += popc4m(lo, M41);
        l31 += popc4m(lo, M31);
      }
    }
    if (MODE == kModeCount) continue;
    uint32_t step_last;
    const uint32_t prev0 = max(prev_in_step(chunk_last(q, boff), lane, step_last), carry);
    carry = max(carry, step_last);
    const uint32_t vb = __umul24(240u, boff);
    // np = vb - 8 * (the last prime so far), or 0 before the wave's first
    uint32_t np = prev0 ? vb - 8u * prev0 : 0u, lm = 0;
    const uint32_t ws[4] = {q.x, q.y, q.z, q.w};
    // bit-reversed words: byte 3 - k of rs[i] is byte k of ws[i] reversed, so
    // its lowest set bit is 7 minus the highest set bit of byte k
    const uint32_t rs[4] = {__builtin_bitreverse32(q.x), __builtin_bitreverse32(q.y), __builtin_bitreverse32(q.z),
                            __builtin_bitreverse32(q.w)};
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const uint32_t x = (ws[j >> 2] >> (8 * (j & 3))) & 0xFFu;
      // 8 * W30 of the first and last set bit from two byte tables (v_perm).
      // The selector's upper bytes are 0, so the result's upper bytes repeat
      // table byte 0: 0x08080800 + 8 * W30 (first) and 0xE8E8E800 + 8 * W30
      // (last); the offsets are folded into the constants below.  A zero
      // byte has t = 0xFFFFFFFF, every selector byte reads 0xFF, pf = -1 and
      // g8 = np + 240 j, its distance from the last prime (times 8); a
      // non-zero byte has g8 = D + 8 * (its gap).
      const uint32_t t = ffbl_byte(ws[j >> 2], j);
      const uint32_t pf = __builtin_amdgcn_perm(kW30x8Hi, kW30x8Lo, t);
      const uint32_t g8 = pf + np + (240u * j + 1u);
      lm = max(lm, g8);
      if (x) {
        lds_add((lds_u32*)(uintptr_t)(H8m + 2u * min(g8, kCap8 + D)), 1u);
        const uint32_t pl = __builtin_amdgcn_perm(kW30x8RevHi, kW30x8RevLo, ffbl_byte(rs[j >> 2], 3 - (j & 3)));
        np = (0xE8E8E800u - 240u * j) - pl;
      }
      // two or more set bits: x is not its lowest set bit (1 << t; 0 for x = 0)
      if (x != ((1u << (t & 31)) & 0xFFu)) lds_add((lds_u32*)((lds_u8*)(uintptr_t)(4 * x) + kLdsByteHist * 4), 1u);
    }
    if (lm >= D) lm -= D;
    if (prev0 == 0 && (q.x | q.y | q.z | q.w) != 0) {
      const uint32_t vfirst = chunk_first(q, boff);
      lds_add((lds_u32*)(H8 + 2 * min(8u * vfirst - vb, kCap8)), 0xFFFFFFFFu);
      wfirst = vfirst;
    }
    lm >>= 3;
    // a chunk wholly past the last counted value has no prime to close its
    // zero run: its distance from the last prime bounds no gap
    if (partial && 8 * boff >= g.vhi) lm = 0;
    // internal gaps are at most 28; count them in the maximum only coarsely
    // (an over-estimate only triggers a records rescan of tiny segments)
    maxg = max(maxg, lm);
    if (((q.x | q.y | q.z | q.w) != 0) && lm < 28) maxg = max(maxg, 28u);
    if (lm >= slow) chunk_big_gaps<MODE>(A, lds, q, boff, prev0, g, copy);
  }
  __builtin_amdgcn_s_setprio(0);
  // per-wave reductions
  cnt = wave_sum(cnt);
  if (MODE == kModeStats) {
    c41 = wave_sum(c41);
    c31 = wave_sum(c31);
    lcnt = wave_sum(lcnt);
    l41 = wave_sum(l41);
    l31 = wave_sum(l31);
  }
  if (MODE != kModeCount) {
    maxg = wave_max(maxg);
    wfirst = wave_max(wfirst);
  }
  if (lane == 0) {
    uint32_t* m = lds + kLdsMisc + kMiscWave + 8 * wave;
    m[0] = wfirst; m[1] = carry; m[2] = cnt; m[3] = c41; m[4] = c31; m[5] = maxg;
    m[6] = lcnt; m[7] = l41;
    lds[kLdsMisc + kMiscWave + 8 * NW + wave] = l31;
  }
}

// Flush the LDS histograms: byte values are first folded into the gap bins
// through their internal gaps (LDS atomics), then each non-zero gap bin goes
// to sink(bin, count) once, so a flush costs at most kHistBins global
// atomics.  Called by every thread of the workgroup (it has barriers).
template <int BS, typename Sink>
__device__ void flush_histograms(uint32_t* lds, Sink sink) {
  for (uint32_t x = threadIdx.x; x < 256; x += BS) {
    const uint32_t s = lds[kLdsByteHist + x];
    lds[kLdsByteHist + x] = 0;
    if (!s) continue;
    for (uint32_t nib = internal_gaps(x); nib; nib >>= 4)
      lds_add(lds + kLdsHist + (nib & 15) * kHistCopies + (x & (kHistCopies - 1)), s);
  }
  __syncthreads();
  for (uint32_t h = threadIdx.x; h < kHistBins; h += BS) {
    uint32_t s = 0;
#pragma unroll
    for (uint32_t c = 0; c < kHistCopies; ++c) {
      s += lds[kLdsHist + h * kHistCopies + c];
      lds[kLdsHist + h * kHistCopies + c] = 0;
    }
    if (s) sink(h, s);
  }
}

template <int BS>
__device__ void flush_hist_stats(const SieveArgs& A, uint32_t* lds) {
  flush_histograms<BS>(lds, [&](uint32_t h, uint32_t n) { atomicAdd(A.hist + h, (unsigned long long)n); });
}

template <int BS>
__device__ void flush_hist_query(const SieveArgs& A, uint32_t* lds, uint32_t qi) {
  unsigned int* qh = A.qhist + (size_t)qi * kQueryBins;
  flush_histograms<BS>(lds, [&](uint32_t h, uint32_t n) { atomicAdd(qh + ((h >= 1 && h <= 256) ? h : 0u), n); });
}

__device__ __forceinline__ SegGeom geometry(const SieveArgs& A, uint32_t s, int mode) {
  SegGeom g;
  if (mode == kModeBatch) {
    const uint32_t qi = A.seg_query[s];
    const QueryDesc Q = A.queries[qi];
    g.query = qi;
    g.k0 = Q.k_lo + (uint64_t)(s - Q.seg_first) * A.seg_bytes;
    uint64_t rem = Q.k_hi - g.k0;
    g.nbytes = rem < A.seg_bytes ? (uint32_t)rem : A.seg_bytes;
    g.lo_val = Q.a;
    g.hi_val = Q.b;
    g.from_cache = Q.from_cache;
  } else {
    g.query = 0;
    g.from_cache = 0;
    g.k0 = A.k_base + (uint64_t)(A.seg_base + s) * A.seg_bytes;
    uint64_t rem = A.k_end - g.k0;
    g.nbytes = rem < A.seg_bytes ? (uint32_t)rem : A.seg_bytes;
    g.lo_val = A.lo_val;
    g.hi_val = A.hi_val;
  }
  g.vlo = bit_at_or_above(g.lo_val, g.k0, g.nbytes);
  g.vhi = bit_at_or_above(g.hi_val + 1, g.k0, g.nbytes);
  return g;
}

}  // namespace

// ---------------------------------------------------------------- main kernel
template <int MODE, int BS>
__global__ void __launch_bounds__(BS) sieve_kernel(SieveArgs A) {
  extern __shared__ uint32_t lds[];
  constexpr int NW = BS / 64;
  init_tables(lds);
  for (uint32_t i = threadIdx.x; i < kHistBins * kHistCopies + 256; i += BS) lds[kLdsHist + i] = 0;
  uint32_t cur_query = 0xFFFFFFFFu;
  uint32_t since_flush = 0;
  __syncthreads();
  // Workgroup w starts on segment w.  Sweeps (count and stats) then take
  // their next segment from the work counter (zeroed by the host): thread 64
  // draws it while the current segment is sieved and every thread reads it
  // at the loop latch, after the walk's barrier (the slot is rewritten only
  // after the next init barrier, which wave 0 reaches after its merge), so
  // workgroups that drew cheaper segments take more and all finish within
  // one segment of each other (static round-robin: 1e11 sweep +1.1 %, 1e12
  // +1 %).  Bitmap mode keeps round-robin.  The loop bound and the segment
  // index are wave-uniform, and every thread-0 region below is followed by a
  // barrier in uniform code (a divergent region at the loop latch lets the
  // compiler hoist it into a divergent inner loop around the next barrier).
  const uint32_t nseg = (MODE == kModeBatch && A.nseg_dev) ? *A.nseg_dev : A.nseg;
  uint32_t s = blockIdx.x;
  if (MODE == kModeBatch) {
    // dynamic work-list: thread 0 takes the next segment; the item is read
    // after one barrier and the slot reused only after the end-of-iteration
    // barrier, so the loop stays wave-uniform
    if (threadIdx.x == 0) lds[kLdsMisc + kMiscItem] = atomicAdd(A.work, 1u);
    __syncthreads();
    s = lds[kLdsMisc + kMiscItem];
  }
  while (s < nseg) {
    const SegGeom g = geometry(A, s, MODE);
    if (MODE == kModeBatch && g.query != cur_query) {
      if (cur_query != 0xFFFFFFFFu) flush_hist_query<BS>(A, lds, cur_query);
      cur_query = g.query;
    }
    if (MODE == kModeStats && threadIdx.x == 0) {
      // checkpoint split: at most one bucket boundary per segment; computed
      // once here (two 64-bit divisions) and read by every thread after the
      // barriers below
      uint32_t split = 0xFFFFFFFFu, b0 = 0, b1 = 0;
      if (A.cp_interval) {
        const uint64_t vfirst = 30 * g.k0 + 1;
        const uint64_t vlast = 30 * (g.k0 + g.nbytes) - 1;
        b0 = (uint32_t)((vfirst - 1) / A.cp_interval);
        b1 = (uint32_t)((vlast - 1) / A.cp_interval);
        if (b1 != b0) split = bit_at_or_above((uint64_t)(b0 + 1) * A.cp_interval + 1, g.k0, g.nbytes);
      }
      lds[kLdsMisc + kMiscSplit] = split;
      lds[kLdsMisc + kMiscSplit + 1] = b0;
      lds[kLdsMisc + kMiscSplit + 2] = b1;
    }
    const bool cached = MODE == kModeBatch && g.from_cache;  // uniform per workgroup
    if (cached) {
      cache_segment<BS>(A, lds, g);
    } else {
      init_segment<BS>(A, lds, g);
    }
    __syncthreads();
    if ((MODE == kModeStats || MODE == kModeCount) && threadIdx.x == 64)
      lds[kLdsMisc + kMiscItem] = gridDim.x + atomicAdd(A.work, 1u);
    if (!cached) sieve_segment<BS>(A, lds, g);
    // the walk's first read-pattern words (all ones where none apply), loaded
    // before the barrier so that their latency overlaps the wait for the
    // other waves (the barrier waits for LDS operations only)
    PatternWords bfirst;
#pragma unroll
    for (int i = 0; i < kReadPatterns; ++i) bfirst.w[i] = make_uint4(~0u, ~0u, ~0u, ~0u);
    if (MODE != kModeBitmap && !cached) {
      const uint32_t wbeg = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) * (A.seg_bytes / NW);
      if (wbeg < g.nbytes) bfirst = load_pattern_words(read_patterns(A, g.k0), wbeg + 16 * (threadIdx.x & 63));
    }
    __syncthreads();
    if (MODE == kModeBitmap) {
      const uint8_t* seg8 = reinterpret_cast<const uint8_t*>(lds + kLdsSeg);
      uint8_t* dst = A.bitmap + (g.k0 - A.k_base);
      const uint32_t n16 = (g.nbytes + 15) >> 4;
      const ReadPatterns RP = read_patterns(A, g.k0);
      for (uint32_t t = threadIdx.x; t < n16; t += BS) {
        uint4 q = *reinterpret_cast<const uint4*>(seg8 + 16 * t);
        and4(q, read_pattern_chunk(RP, 16 * t));
        const uint32_t cb = 128 * t;
        if (cb < g.vlo || cb + 128 > g.vhi) mask_chunk(q, cb, g.vlo, g.vhi);
        *reinterpret_cast<uint4*>(dst + 16 * t) = q;
      }
      __syncthreads();
      s += gridDim.x;  // bitmap mode is a sweep: static round-robin
      continue;
    }

    uint32_t split = 0xFFFFFFFFu, b0 = 0, b1 = 0;
    if (MODE == kModeStats) {
      split = lds[kLdsMisc + kMiscSplit];
      b0 = lds[kLdsMisc + kMiscSplit + 1];
      b1 = lds[kLdsMisc + kMiscSplit + 2];
    }
    walk_segment<MODE, BS>(A, lds, g, split, bfirst);
    __syncthreads();

    if (threadIdx.x < 64) {
      // Wave 0 joins the per-wave records, lane w holding wave w's: the
      // waves' last primes increase with w, so a max-scan gives each wave
      // the last prime before it; sums and maxima are 16-lane reductions
      // (lanes >= NW hold zeros) and the global adds go out one per lane.
      const uint32_t lane = threadIdx.x;
      const uint32_t* mw = lds + kLdsMisc + kMiscWave + 8 * (lane < NW ? lane : 0);
      const bool mine = lane < NW;
      const uint32_t wc = mine ? mw[2] : 0u;
      const bool has = wc != 0;
      const uint32_t wf = has ? mw[0] : 0u, wl = has ? mw[1] : 0u;
      uint32_t incl = wl;
#pragma unroll
      for (int d = 1; d < NW; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= (uint32_t)d) incl = max(incl, t);
      }
      uint32_t before = (uint32_t)__shfl_up((int)incl, 1, 64);
      if (lane == 0) before = 0;
      uint32_t smax = 0;
      if (MODE != kModeCount && has) {
        smax = mw[5];
        if (before) note_gap<MODE>(A, lds, wf - before, wf, g, 0, smax);
      }
      if (MODE == kModeStats && A.wave_gaps && mine) {
        // for the records rescan: the wave's largest gap (the gap into its
        // first prime included) and the distance from the segment's last
        // earlier prime to the wave's first value.  Both are below one gap
        // between primes (< 1600 below 2^64), so 16 bits hold each.
        const uint32_t d = before ? 30u * lane * (A.seg_bytes / NW) - before : 0u;
        A.wave_gaps[(size_t)(A.seg_base + s) * NW + lane] = (min(d, 0xFFFFu) << 16) | min(smax, 0xFFFFu);
      }
      const uint64_t hb = __ballot(has);
      const uint32_t sfirst = (uint32_t)__shfl((int)wf, hb ? (int)__builtin_ctzll(hb) : 0, 64);
      const uint32_t slast = (uint32_t)__shfl((int)incl, NW - 1, 64);
      auto sum16 = [](uint32_t v) {
#pragma unroll
        for (int d = 8; d >= 1; d >>= 1) v += (uint32_t)__shfl_xor((int)v, d, 16);
        return v;
      };
      const uint32_t scnt = sum16(wc);
      if (MODE != kModeCount) {
#pragma unroll
        for (int d = 8; d >= 1; d >>= 1) smax = max(smax, (uint32_t)__shfl_xor((int)smax, d, 16));
      }
      if (lane == 0) {
        if (A.summaries) {
          SegSummary ss;
          ss.first = sfirst; ss.last = slast; ss.maxgap = smax; ss.count = scnt;
          A.summaries[(MODE == kModeBatch) ? s : A.seg_base + s] = ss;
        }
        if (MODE == kModeBatch) atomicAdd(A.qcount + g.query, (unsigned long long)scnt);
        if (MODE == kModeCount) atomicAdd(A.totals + 0, (unsigned long long)scnt);
      }
      if (MODE == kModeStats) {
        const uint32_t s41 = sum16(mine ? mw[3] : 0u), s31 = sum16(mine ? mw[4] : 0u);
        const uint32_t lc = sum16(mine ? mw[6] : 0u), l41 = sum16(mine ? mw[7] : 0u);
        const uint32_t l31 = sum16(mine ? lds[kLdsMisc + kMiscWave + 8 * NW + lane] : 0u);
        // lanes 0-2: totals; 3-5: the first checkpoint bucket; 6-8: the
        // second (segments split by a checkpoint)
        const uint32_t k = lane % 3;
        const uint32_t vf = k == 0 ? scnt : (k == 1 ? s41 : s31);
        const uint32_t vl = k == 0 ? lc : (k == 1 ? l41 : l31);
        unsigned long long* dst = nullptr;
        uint32_t val = 0;
        if (lane < 3) {
          dst = A.totals + k;
          val = vf;
        } else if (lane < 6) {
          if (A.cp_interval && b0 < A.nbuckets) {
            dst = A.buckets + 3 * b0 + k;
            val = split == 0xFFFFFFFFu ? vf : vl;
          }
        } else if (lane < 9) {
          if (A.cp_interval && split != 0xFFFFFFFFu && b1 < A.nbuckets) {
            dst = A.buckets + 3 * b1 + k;
            val = vf - vl;
          }
        }
        if (dst) atomicAdd(dst, (unsigned long long)val);
      }
    }
    // Sweeps need no barrier here: the next segment's init touches nothing
    // the merge reads, so waves 1.. start it (and its pattern copies, which
    // wave 0 leaves to them) while wave 0 merges, and the init barrier joins
    // them.  The histogram flush reads what the merge adds: it waits first.
    if (MODE == kModeBatch) __syncthreads();
    if (MODE == kModeStats && ++since_flush == 1024) {
      __syncthreads();
      flush_hist_stats<BS>(A, lds);
      since_flush = 0;
      __syncthreads();
    }
    if (MODE == kModeBatch) {
      if (threadIdx.x == 0) lds[kLdsMisc + kMiscItem] = atomicAdd(A.work, 1u);
      __syncthreads();
      s = lds[kLdsMisc + kMiscItem];
    } else if (MODE == kModeStats || MODE == kModeCount) {
      s = __builtin_amdgcn_readfirstlane(lds[kLdsMisc + kMiscItem]);
    } else {
      s += gridDim.x;
    }
  }
  if (MODE == kModeStats) {
    __syncthreads();
    flush_hist_stats<BS>(A, lds);
  }
  if (MODE == kModeBatch && cur_query != 0xFFFFFFFFu) flush_hist_query<BS>(A, lds, cur_query);
}

// ---------------------------------------------------------------- stitch (sweep)
// One workgroup joins the per-segment summaries in order: boundary gaps go
// to the histogram and large-gap list; boundary records are emitted
// directly; within segments whose internal largest gap beats every earlier
// gap, each wave range (the bytes one wave walked) that raises the running
// maximum is listed for rescan_kernel with the maximum before it.
// Exclusive scan of one value per thread over a 1024-thread block, in thread
// order, under an associative op(earlier, later) with identity `ident`:
// 64-lane shuffle scans, then a scan of the 16 wave totals.  Values are
// 64-bit (two 32-bit shuffles each step).  `tot` is 32 words of shared
// scratch; every thread must call it (it has barriers).
template <typename Op>
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t ident, Op op, unsigned long long* tot) {
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  auto up = [](uint64_t x, int d) -> uint64_t {
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)x, d, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(x >> 32), d, 64);
    return ((uint64_t)hi << 32) | lo;
  };
  uint64_t incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint64_t u = up(incl, d);
    if (lane >= (uint32_t)d) incl = op(u, incl);
  }
  uint64_t excl = up(incl, 1);
  if (lane == 0) excl = ident;
  if (lane == 63) tot[wave] = incl;
  __syncthreads();
  if (wave == 0) {
    uint64_t w = lane < 16 ? (uint64_t)tot[lane] : ident;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const uint64_t u = up(w, d);
      if (lane >= (uint32_t)d) w = op(u, w);
    }
    if (lane < 16) tot[16 + lane] = w;
  }
  __syncthreads();
  const uint64_t before = wave ? (uint64_t)tot[16 + wave - 1] : ident;
  return op(before, excl);
}

__global__ void __launch_bounds__(1024) stitch_kernel(StitchArgs A) {
  __shared__ unsigned long long s_tot[32];
  __shared__ uint32_t s_hist[kHistBins];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < kHistBins; i += 1024) s_hist[i] = 0;
  // records cluster in the first segments (maximal gaps grow slowly), and a
  // segment that holds one costs a thread a chain of loads and atomics: the
  // first 64 segments go one to a thread (threads 0-63), the rest evenly to
  // threads 64-1023
  const uint32_t early = min(A.nseg, 64u);
  const uint32_t per = (A.nseg - early + 959) / 960;
  const uint32_t lo = t < 64 ? min(t, early) : min(early + (t - 64) * per, A.nseg);
  const uint32_t hi = t < 64 ? min(t + 1, early) : min(lo + per, A.nseg);
  auto absval = [&](uint32_t s, uint32_t rel) -> uint64_t {
    return 30ull * (A.k_base + (uint64_t)s * A.seg_bytes) + rel;
  };
  // the thread's summaries in groups of 8, every load of a group issued
  // before the first is used (one memory latency per group, not per segment)
  auto for_each_summary = [&](auto&& f) {
    for (uint32_t s0 = lo; s0 < hi; s0 += 8) {
      SegSummary b[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) b[j] = A.summaries[min(s0 + j, hi - 1)];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (s0 + j < hi) f(s0 + j, b[j]);
    }
  };
  uint64_t last = 0, ffirst = 0;
  uint32_t emax = 0;
  for_each_summary([&](uint32_t s, const SegSummary& S) {
    emax = max(emax, S.maxgap);
    if (!S.count) return;
    const uint64_t F = absval(s, S.first);
    if (last) emax = max(emax, (uint32_t)(F - last));
    else ffirst = F;
    last = absval(s, S.last);
  });
  // the last prime before this thread's segments: the last non-zero "last"
  // of the threads before it, else the seed
  const uint64_t lastb = block_exclusive_scan(last, 0, [](uint64_t a, uint64_t b) { return b ? b : a; }, s_tot);
  const uint64_t prev0 = lastb ? lastb : (uint64_t)A.seed_last;
  // boundary gap of this thread's first non-empty segment
  if (ffirst) emax = max(emax, (uint32_t)(ffirst - prev0));
  __syncthreads();  // s_tot is reused
  // the largest gap before this thread's segments
  uint32_t run = (uint32_t)block_exclusive_scan(emax, 0, [](uint64_t a, uint64_t b) { return a > b ? a : b; }, s_tot);
  run = max(run, (uint32_t)A.seed_max);
  uint64_t prev = prev0;
  for_each_summary([&](uint32_t s, const SegSummary& S) {
    if (!S.count) return;
    const uint64_t F = absval(s, S.first);
    const uint32_t B = (uint32_t)(F - prev);
    const uint32_t h = B >> 1;
    if (h < kHistBins) lds_add(s_hist + h, 1u);
    else atomicAdd(A.hist + (h < kGapBins ? h : kGapBins - 1), 1ull);
    if (B >= A.large_threshold) {
      unsigned int slot = atomicAdd(A.n_large, 1u);
      if (slot < A.large_cap) { A.large[slot].position = F; A.large[slot].gap = B; A.large[slot].seg = s; }
    }
    if (B > run) {
      unsigned int slot = atomicAdd(A.n_rec, 1u);
      if (slot < A.rec_cap) { A.rec[slot].position = F; A.rec[slot].gap = B; A.rec[slot].seg = s; }
      run = B;
    }
    if (S.maxgap > run) {
      // list each wave range of the segment that raises the running maximum:
      // the segment's wave words in four 16-byte loads (one latency), the
      // ranges counted, then one slot reservation for all of them
      const uint4* wg = reinterpret_cast<const uint4*>(A.wave_gaps + (size_t)s * A.waves);
      uint32_t x[16];
#pragma unroll
      for (uint32_t w4 = 0; w4 < 4; ++w4) {
        const uint4 v = 4 * w4 < A.waves ? wg[w4] : make_uint4(0u, 0u, 0u, 0u);
        x[4 * w4] = v.x; x[4 * w4 + 1] = v.y; x[4 * w4 + 2] = v.z; x[4 * w4 + 3] = v.w;
      }
      uint32_t r = run, n = 0;
#pragma unroll
      for (uint32_t w = 0; w < 16; ++w)
        if ((x[w] & 0xFFFFu) > r) { ++n; r = x[w] & 0xFFFFu; }
      if (n) {
        const unsigned int slot = atomicAdd(A.n_rescan, n);
        uint32_t i = 0;
        r = run;
#pragma unroll
        for (uint32_t w = 0; w < 16; ++w)
          if ((x[w] & 0xFFFFu) > r) {
            if (slot + i < A.rescan_cap) A.rescan[slot + i] = RescanItem{s, r, w, x[w] >> 16};
            ++i;
            r = x[w] & 0xFFFFu;
          }
      }
      run = max(r, S.maxgap);
    }
    prev = absval(s, S.last);
  });
  if (t == 1023) {
    A.final_out[0] = prev;
    A.final_out[1] = run;
  }
  __syncthreads();
  for (uint32_t i = t; i < kHistBins; i += 1024)
    if (s_hist[i]) atomicAdd(A.hist + i, (unsigned long long)s_hist[i]);
}

// ---------------------------------------------------------------- records rescan
// One workgroup per flagged wave range (the seg_bytes / waves bytes that one
// wave of the sweep walked): re-sieve just that range and list every gap
// larger than all earlier gaps, starting from the running maximum before the
// range and from the last prime before it.  The range is walked in 1 KiB
// steps, a wave per step: pass 1 takes each step's first and last prime and
// largest inner gap, thread 0 chains the steps in order, and pass 2 re-walks
// only the steps that hold a record, each lane emitting the gaps of its chunk
// above the larger of the step's prior maximum and the earlier lanes' gaps.

// Largest gap in a lane's 16-byte chunk at segment-relative byte boff (prev =
// last prime before it, 0 if none), optionally emitting the gaps above `run`
// in order.
template <bool EMIT>
__device__ __forceinline__ uint32_t chunk_gaps(uint4 q, uint32_t boff, uint32_t prev, uint32_t run,
                                               const RescanArgs& R, uint64_t k0, uint32_t seg) {
  uint32_t lm = 0;
  const uint32_t ws[4] = {q.x, q.y, q.z, q.w};
  for (int wi = 0; wi < 4; ++wi) {
    uint32_t w = ws[wi];
    while (w) {
      const uint32_t b = __builtin_ctz(w);
      w &= w - 1;
      const uint32_t v = 30 * (boff + 4 * wi + (b >> 3)) + w30(b & 7);
      if (prev) {
        const uint32_t gap = v - prev;
        lm = max(lm, gap);
        if (EMIT && gap > run) {
          run = gap;
          const unsigned int slot = atomicAdd(R.n_out, 1u);
          if (slot < R.out_cap) {
            R.out[slot].position = 30 * k0 + v;
            R.out[slot].gap = gap;
            R.out[slot].seg = seg;
          }
        }
      }
      prev = v;
    }
  }
  return lm;
}

// 1 KiB steps of the largest wave range (seg_bytes / 8 waves of a 512-thread
// block): five per-step arrays of this length fill misc words 16..115
constexpr uint32_t kMaxRescanSteps = 20;
static_assert(kMaxSegBytes / 8 <= kMaxRescanSteps * 1024, "rescan step records");
static_assert(16 + 5 * kMaxRescanSteps <= kMiscBounds, "rescan step records below the segment bounds");

template <int BS>
__global__ void __launch_bounds__(BS) rescan_kernel(SieveArgs A, RescanArgs R) {
  extern __shared__ uint32_t lds[];
  constexpr int NW = BS / 64;
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  init_tables(lds);
  __syncthreads();
  const uint32_t nitems = min(*R.n_items, R.items_cap);  // same value in every thread
  const uint32_t sub = A.seg_bytes / NW;                  // bytes of one wave range (a multiple of 1 KiB)
  for (uint32_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const RescanItem item = R.items[it];
    const uint32_t s = item.seg, off = item.wave * sub;
    const uint64_t kseg = A.k_base + (uint64_t)s * A.seg_bytes;
    SegGeom g;
    g.query = 0;
    g.from_cache = 0;
    g.k0 = kseg + off;
    const uint64_t rem = A.k_end - g.k0;  // > 0: the range holds primes
    g.nbytes = rem < sub ? (uint32_t)rem : sub;
    g.lo_val = A.lo_val;
    g.hi_val = A.hi_val;
    g.vlo = bit_at_or_above(g.lo_val, g.k0, g.nbytes);
    g.vhi = bit_at_or_above(g.hi_val + 1, g.k0, g.nbytes);
    init_segment<BS>(A, lds, g);
    __syncthreads();
    sieve_segment<BS>(A, lds, g);
    __syncthreads();

    // Pass 1, a wave per 1 KiB step: the step's first and last prime and
    // its largest gap between two of its own primes.  Values are
    // segment-relative, as in the sweep.  Steps read below ceil(nbytes / 1 KiB)
    // KiB <= sub bytes of the LDS segment; bits past nbytes are masked.
    const uint8_t* seg8 = reinterpret_cast<const uint8_t*>(lds + kLdsSeg);
    const uint32_t nsteps = (g.nbytes + 1023) >> 10;  // <= kMaxRescanSteps
    uint32_t* st_first = lds + kLdsMisc + 16;  // misc words 16..115: per-step records
    uint32_t* st_last = st_first + kMaxRescanSteps;
    uint32_t* st_max = st_last + kMaxRescanSteps;
    uint32_t* st_thr = st_max + kMaxRescanSteps;    // largest gap before the step
    uint32_t* st_carry = st_thr + kMaxRescanSteps;  // last prime before the step
    const ReadPatterns RP = read_patterns(A, g.k0);
    auto load_chunk = [&](uint32_t st) {
      const uint32_t loc = 1024 * st + 16 * lane;
      uint4 q = *reinterpret_cast<const uint4*>(seg8 + loc);
      and4(q, read_pattern_chunk(RP, loc));
      const uint32_t cb = 8 * loc;
      if (cb < g.vlo || cb + 128 > g.vhi) mask_chunk(q, cb, g.vlo, g.vhi);
      return q;
    };
    for (uint32_t st = wave; st < nsteps; st += NW) {
      const uint4 q = load_chunk(st);
      const uint32_t boff = off + 1024 * st + 16 * lane;
      uint32_t step_last;
      const uint32_t prev = prev_in_step(chunk_last(q, boff), lane, step_last);
      const uint32_t lm = wave_max(chunk_gaps<false>(q, boff, prev, 0, R, kseg, s));
      const uint32_t cf = chunk_first(q, boff);
      const uint64_t with = __ballot(cf != 0);
      const uint32_t sf = with ? (uint32_t)__shfl((int)cf, (int)__builtin_ctzll(with), 64) : 0u;
      if (lane == 0) {
        st_first[st] = sf;
        st_last[st] = step_last;
        st_max[st] = lm;
      }
    }
    __syncthreads();
    // the steps in order: the running maximum and the last prime before each
    if (threadIdx.x == 0) {
      uint32_t run = item.threshold;
      uint32_t carry = item.dprev ? 30 * off - item.dprev : 0;
      for (uint32_t st = 0; st < nsteps; ++st) {
        st_thr[st] = run;
        st_carry[st] = carry;
        const uint32_t f = st_first[st];
        if (!f) continue;
        if (carry) run = max(run, f - carry);
        run = max(run, st_max[st]);
        carry = st_last[st];
      }
    }
    __syncthreads();
    // Pass 2, only for steps holding a gap above the maximum before them:
    // each lane emits its chunk's gaps above the larger of that maximum and
    // the earlier lanes' largest gap.
    for (uint32_t st = wave; st < nsteps; st += NW) {
      const uint32_t f = st_first[st], carry = st_carry[st], thr = st_thr[st];
      if (!f || max(st_max[st], carry ? f - carry : 0u) <= thr) continue;  // wave-uniform
      const uint4 q = load_chunk(st);
      const uint32_t boff = off + 1024 * st + 16 * lane;
      uint32_t step_last;
      const uint32_t prev = max(prev_in_step(chunk_last(q, boff), lane, step_last), carry);
      const uint32_t lm = chunk_gaps<false>(q, boff, prev, 0, R, kseg, s);
      uint32_t incl = lm;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = (uint32_t)__shfl_up((int)incl, d, 64);
        if (lane >= (uint32_t)d) incl = max(incl, t);
      }
      uint32_t excl = (uint32_t)__shfl_up((int)incl, 1, 64);
      if (lane == 0) excl = 0;
      const uint32_t r = max(thr, excl);
      if (lm > r) chunk_gaps<true>(q, boff, prev, r, R, kseg, s);
    }
    __syncthreads();  // the next item rebuilds the range in LDS
  }
}

// ---------------------------------------------------------------- select
// One workgroup per request: sieve sweep segment segs[i] (byte k0 = seg *
// seg_bytes, values >= 7 only) and return its ranks[i]-th prime (1-based),
// or 0 when the segment holds fewer.
template <int BS>
__global__ void __launch_bounds__(BS) select_kernel(SieveArgs A, const uint32_t* segs, const uint32_t* ranks,
                                                    unsigned long long* out, const uint32_t* n_dev) {
  extern __shared__ uint32_t lds[];
  constexpr int NW = BS / 64;
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (blockIdx.x >= *n_dev) return;  // uniform per workgroup
  init_tables(lds);
  __syncthreads();
  const uint32_t s = segs[blockIdx.x];
  const uint32_t rank = ranks[blockIdx.x];
  SegGeom g;
  g.query = 0;
  g.from_cache = 0;
  g.k0 = (uint64_t)s * A.seg_bytes;
  const uint64_t rem = A.k_end - g.k0;
  g.nbytes = rem < A.seg_bytes ? (uint32_t)rem : A.seg_bytes;
  g.lo_val = 7;
  g.hi_val = A.hi_val;
  g.vlo = bit_at_or_above(g.lo_val, g.k0, g.nbytes);
  g.vhi = bit_at_or_above(g.hi_val + 1, g.k0, g.nbytes);
  init_segment<BS>(A, lds, g);
  __syncthreads();
  sieve_segment<BS>(A, lds, g);
  __syncthreads();
  const uint8_t* seg8 = reinterpret_cast<const uint8_t*>(lds + kLdsSeg);
  const uint32_t wave_bytes = A.seg_bytes / NW;
  const uint32_t wbeg = wave * wave_bytes, wend = min(wbeg + wave_bytes, g.nbytes);
  uint32_t* wcount = lds + kLdsMisc + 16;  // NW words
  const ReadPatterns RP = read_patterns(A, g.k0);
  uint32_t c = 0;
  for (uint32_t base = wbeg; base < wend; base += 1024) {
    uint4 q = *reinterpret_cast<const uint4*>(seg8 + base + 16 * lane);
    and4(q, read_pattern_chunk(RP, base + 16 * lane));
    const uint32_t cb = 8 * (base + 16 * lane);
    if (cb < g.vlo || cb + 128 > g.vhi) mask_chunk(q, cb, g.vlo, g.vhi);
    c += popc4(q);
  }
  c = wave_sum(c);
  if (lane == 0) wcount[wave] = c;
  __syncthreads();
  // the wave holding the rank-th prime, and the rank within it
  uint32_t before = 0, target_wave = NW;
  for (int w = 0; w < NW; ++w) {
    const uint32_t cw = wcount[w];
    if (target_wave == NW && before + cw >= rank) target_wave = w;
    else if (target_wave == NW) before += cw;
  }
  if (target_wave == NW || rank == 0) {
    if (threadIdx.x == 0) out[blockIdx.x] = 0;
    return;
  }
  if (wave != target_wave) return;
  uint32_t need = rank - before;  // 1-based within this wave's region
  for (uint32_t base = wbeg; base < wend; base += 1024) {
    uint4 q = *reinterpret_cast<const uint4*>(seg8 + base + 16 * lane);
    and4(q, read_pattern_chunk(RP, base + 16 * lane));
    const uint32_t boff = base + 16 * lane;
    const uint32_t cb = 8 * boff;
    if (cb < g.vlo || cb + 128 > g.vhi) mask_chunk(q, cb, g.vlo, g.vhi);
    const uint32_t pc = popc4(q);
    uint32_t incl = pc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = (uint32_t)__shfl_up((int)incl, d, 64);
      if (lane >= (uint32_t)d) incl += t;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
    if (need <= total) {
      const uint32_t excl = incl - pc;
      if (need > excl && need <= incl) {
        uint32_t k = need - excl;  // k-th set bit of this lane's chunk
        const uint32_t ws[4] = {q.x, q.y, q.z, q.w};
        for (int wi = 0; wi < 4; ++wi) {
          uint32_t w = ws[wi];
          const uint32_t pw = __builtin_popcount(w);
          if (k > pw) { k -= pw; continue; }
          for (uint32_t t = 1; t < k; ++t) w &= w - 1;
          const uint32_t b = __builtin_ctz(w);
          out[blockIdx.x] = 30ull * (g.k0 + boff + 4 * wi + (b >> 3)) + w30(b & 7);
          break;
        }
      }
      break;  // wave-uniform
    }
    need -= total;
  }
}

hipError_t launch_select(int block, const SieveArgs& A, const uint32_t* segs, const uint32_t* ranks,
                         unsigned long long* out, const uint32_t* n_dev, uint32_t n_cap, hipStream_t st) {
  if (!n_cap) return hipSuccess;
  if (block == 1024)
    hipLaunchKernelGGL((select_kernel<1024>), dim3(n_cap), dim3(1024), lds_bytes_for(A.seg_bytes), st, A, segs, ranks,
                       out, n_dev);
  else
    hipLaunchKernelGGL((select_kernel<512>), dim3(n_cap), dim3(512), lds_bytes_for(A.seg_bytes), st, A, segs, ranks,
                       out, n_dev);
  return hipGetLastError();
}

// ---------------------------------------------------------------- batch stitch
// One thread per query: join its segments in order (boundary gaps into the
// query histogram) and report first and last prime.
__global__ void batch_stitch_kernel(BatchStitchArgs A) {
  const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nq = A.nq_dev ? *A.nq_dev : A.nq;
  if (qi >= nq) return;
  const QueryDesc Q = A.queries[qi];
  uint64_t first = 0, last = 0;
  unsigned int* qh = A.qhist + (size_t)qi * kQueryBins;
  for (uint32_t i = 0; i < Q.nseg; ++i) {
    const uint32_t s = Q.seg_first + i;
    const SegSummary S = A.summaries[s];
    if (!S.count) continue;
    const uint64_t base = 30ull * (Q.k_lo + (uint64_t)i * A.seg_bytes);
    const uint64_t F = base + S.first;
    if (last) {
      const uint64_t gap = F - last;
      const uint64_t h = gap >> 1;
      atomicAdd(qh + ((h >= 1 && h <= 256 && !(gap & 1)) ? (uint32_t)h : 0u), 1u);
    } else {
      first = F;
    }
    last = base + S.last;
  }
  A.first[qi] = first;
  A.last[qi] = last;
}

// ---------------------------------------------------------------- launchers
template <int MODE, int BS>
static hipError_t launch_sieve_t(const SieveArgs& A, int grid, hipStream_t st) {
  hipLaunchKernelGGL((sieve_kernel<MODE, BS>), dim3(grid), dim3(BS), lds_bytes_for(A.seg_bytes), st, A);
  return hipGetLastError();
}

hipError_t launch_sieve(int mode, int block, const SieveArgs& A, int grid, hipStream_t st) {
  if (block == 1024) {
    if (mode == kModeCount) return launch_sieve_t<kModeCount, 1024>(A, grid, st);
    if (mode == kModeStats) return launch_sieve_t<kModeStats, 1024>(A, grid, st);
    if (mode == kModeBitmap) return launch_sieve_t<kModeBitmap, 1024>(A, grid, st);
    return launch_sieve_t<kModeBatch, 1024>(A, grid, st);
  }
  if (mode == kModeCount) return launch_sieve_t<kModeCount, 512>(A, grid, st);
  if (mode == kModeStats) return launch_sieve_t<kModeStats, 512>(A, grid, st);
  if (mode == kModeBitmap) return launch_sieve_t<kModeBitmap, 512>(A, grid, st);
  return launch_sieve_t<kModeBatch, 512>(A, grid, st);
}

// ---------------------------------------------------------------- bitmap -> sorted list
// Blocks of kListBlock bitmap bytes; 256 threads x 16 bytes each.
constexpr uint32_t kListBlock = 4096;

__global__ void __launch_bounds__(256) list_count_kernel(const uint8_t* bm, uint64_t nbytes, uint32_t* counts) {
  __shared__ uint32_t red[4];
  const uint64_t off = (uint64_t)blockIdx.x * kListBlock + 16 * threadIdx.x;
  uint32_t c = 0;
  if (off < nbytes) c = popc4(*reinterpret_cast<const uint4*>(bm + off));
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) list_write_kernel(const uint8_t* bm, uint64_t nbytes, uint64_t k_base,
                                                         const unsigned long long* block_off, uint64_t* out) {
  __shared__ uint32_t wsum[4];
  const uint32_t lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const uint64_t off = (uint64_t)blockIdx.x * kListBlock + 16 * threadIdx.x;
  uint4 q = make_uint4(0, 0, 0, 0);
  if (off < nbytes) q = *reinterpret_cast<const uint4*>(bm + off);
  const uint32_t c = popc4(q);
  uint32_t incl = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = (uint32_t)__shfl_up((int)incl, d, 64);
    if (lane >= (uint32_t)d) incl += t;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t w = 0; w < wave; ++w) base += wsum[w];
  uint64_t pos = block_off[blockIdx.x] + base + incl - c;
  const uint32_t ws[4] = {q.x, q.y, q.z, q.w};
  for (int wi = 0; wi < 4; ++wi) {
    uint32_t w = ws[wi];
    while (w) {
      const uint32_t b = __builtin_ctz(w);
      w &= w - 1;
      out[pos++] = 30ull * (k_base + off + 4 * wi + (b >> 3)) + w30(b & 7);
    }
  }
}

hipError_t launch_list_count(const uint8_t* bm, uint64_t nbytes, uint32_t* counts, hipStream_t st) {
  const uint64_t nb = (nbytes + kListBlock - 1) / kListBlock;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(list_count_kernel, dim3((uint32_t)nb), dim3(256), 0, st, bm, nbytes, counts);
  return hipGetLastError();
}

hipError_t launch_list_write(const uint8_t* bm, uint64_t nbytes, uint64_t k_base,
                             const unsigned long long* block_off, uint64_t* out, hipStream_t st) {
  const uint64_t nb = (nbytes + kListBlock - 1) / kListBlock;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(list_write_kernel, dim3((uint32_t)nb), dim3(256), 0, st, bm, nbytes, k_base, block_off, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------- LDS probe
// The LDS rate this device sustains for the sieve's access kinds: one
// 1024-thread workgroup per CU over a 128 KiB LDS array (every index masked
// to it), 8 accesses per loop turn.  KIND bit 0: pseudo-random addresses (a
// per-lane linear congruential sequence) instead of consecutive dwords
// across the lanes (conflict-free); bit 1: ds_read_b32 instead of ds_and_b32.
// Block 0's thread 0 also reads the shader clock and the 100 MHz real-time
// counter at both ends, so the host can state the clock it ran at.
constexpr uint32_t kProbeWords = 32768;
template <int KIND>
__global__ void __launch_bounds__(1024) lds_probe_kernel(uint32_t iters, unsigned long long* clocks,
                                                         uint32_t* sink) {
  extern __shared__ uint32_t lds[];
  for (uint32_t i = threadIdx.x; i < kProbeWords; i += 1024) lds[i] = 0xFFFFFFFFu;
  __syncthreads();
  unsigned long long t0 = 0, r0 = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    t0 = __builtin_amdgcn_s_memtime();
    r0 = __builtin_amdgcn_s_memrealtime();
  }
  uint32_t x = threadIdx.x * 2654435761u + blockIdx.x * 40503u + 1u;
  uint32_t acc = 0;
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) {
      uint32_t idx;
      if (KIND & 1) {
        x = x * 1664525u + 1013904223u;
        idx = x >> 17;  // 15 bits: < kProbeWords
      } else {
        idx = (threadIdx.x + 1024u * k + 8192u * it) & (kProbeWords - 1);
      }
      // volatile LDS reads (dynamic LDS starts at address 0): every read issued, none merged
      if (KIND & 2) acc ^= ((volatile lds_u32*)(uintptr_t)0)[idx];
      else __hip_atomic_fetch_and(lds + idx, ~(1u << k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    clocks[0] = __builtin_amdgcn_s_memtime() - t0;
    clocks[1] = __builtin_amdgcn_s_memrealtime() - r0;
  }
  if (acc == 0x9E3779B9u) sink[0] = acc;  // keeps the reads
}

hipError_t launch_lds_probe(int kind, uint32_t blocks, uint32_t iters, unsigned long long* clocks, uint32_t* sink,
                            hipStream_t st) {
  const uint32_t lds = kProbeWords * 4;
  switch (kind) {
    case 0: hipLaunchKernelGGL((lds_probe_kernel<0>), dim3(blocks), dim3(1024), lds, st, iters, clocks, sink); break;
    case 1: hipLaunchKernelGGL((lds_probe_kernel<1>), dim3(blocks), dim3(1024), lds, st, iters, clocks, sink); break;
    case 2: hipLaunchKernelGGL((lds_probe_kernel<2>), dim3(blocks), dim3(1024), lds, st, iters, clocks, sink); break;
    case 3: hipLaunchKernelGGL((lds_probe_kernel<3>), dim3(blocks), dim3(1024), lds, st, iters, clocks, sink); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t prepare_kernels(uint32_t lds_bytes) {
  const void* fns[] = {
      (const void*)sieve_kernel<kModeCount, 1024>, (const void*)sieve_kernel<kModeStats, 1024>,
      (const void*)sieve_kernel<kModeBatch, 1024>, (const void*)sieve_kernel<kModeCount, 512>,
      (const void*)sieve_kernel<kModeStats, 512>,  (const void*)sieve_kernel<kModeBatch, 512>,
      (const void*)sieve_kernel<kModeBitmap, 1024>, (const void*)sieve_kernel<kModeBitmap, 512>,
      (const void*)rescan_kernel<1024>,            (const void*)rescan_kernel<512>,
      (const void*)select_kernel<1024>,            (const void*)select_kernel<512>,
      (const void*)lds_probe_kernel<0>,            (const void*)lds_probe_kernel<1>,
      (const void*)lds_probe_kernel<2>,            (const void*)lds_probe_kernel<3>};
  for (const void* f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    // the walk addresses LDS tables absolutely: dynamic LDS must start at 0
    hipFuncAttributes fa;
    e = hipFuncGetAttributes(&fa, f);
    if (e != hipSuccess) return e;
    if (fa.sharedSizeBytes != 0) return hipErrorInvalidKernelFile;
  }
  return hipSuccess;
}

hipError_t launch_stitch(const StitchArgs& A, hipStream_t st) {
  hipLaunchKernelGGL(stitch_kernel, dim3(1), dim3(1024), 0, st, A);
  return hipGetLastError();
}

hipError_t launch_rescan(int block, const SieveArgs& A, const RescanArgs& R, int grid, hipStream_t st) {
  if (grid <= 0) return hipSuccess;
  if (block == 1024)
    hipLaunchKernelGGL((rescan_kernel<1024>), dim3(grid), dim3(1024), lds_bytes_for(A.seg_bytes), st, A, R);
  else
    hipLaunchKernelGGL((rescan_kernel<512>), dim3(grid), dim3(512), lds_bytes_for(A.seg_bytes), st, A, R);
  return hipGetLastError();
}

hipError_t launch_batch_stitch(const BatchStitchArgs& A, uint32_t nq_cap, hipStream_t st) {
  if (nq_cap == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_stitch_kernel, dim3((nq_cap + 255) / 256), dim3(256), 0, st, A);
  return hipGetLastError();
}

__global__ void zero_kernel(uint32_t* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

hipError_t launch_zero(uint32_t* p, uint64_t n, hipStream_t st) {
  if (!n) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

// out[0] = add + the counts of segments [0, n), summed by one workgroup.
__global__ void __launch_bounds__(1024) sum_counts_kernel(const SegSummary* sums, uint32_t n, uint64_t add,
                                                           unsigned long long* out) {
  __shared__ unsigned long long part[16];
  unsigned long long c = 0;
  for (uint32_t i = threadIdx.x; i < n; i += 1024) c += sums[i].count;
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = add;
    for (int w = 0; w < 16; ++w) t += part[w];
    out[0] = t;
  }
}

hipError_t launch_sum_counts(const SegSummary* sums, uint32_t n, uint64_t add, unsigned long long* out,
                             hipStream_t st) {
  hipLaunchKernelGGL(sum_counts_kernel, dim3(1), dim3(1024), 0, st, sums, n, add, out);
  return hipGetLastError();
}

__global__ void spans_kernel(SpanList L) {
  const CopySpan S = L.s[blockIdx.y];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S.words; i += (uint64_t)gridDim.x * blockDim.x)
    S.dst[i] = S.src ? S.src[i] : 0u;
}

hipError_t launch_spans(const SpanList& L, hipStream_t st) {
  if (L.n == 0) return hipSuccess;
  if (L.n > (uint32_t)kMaxSpans) return hipErrorInvalidValue;
  uint64_t mx = 1;
  for (uint32_t j = 0; j < L.n; ++j) mx = std::max<uint64_t>(mx, L.s[j].words);
  const uint64_t blocks = std::min<uint64_t>((mx + 255) / 256, 1024);
  hipLaunchKernelGGL(spans_kernel, dim3((uint32_t)blocks, L.n), dim3(256), 0, st, L);
  return hipGetLastError();
}

}  // namespace voss
