// host_math.cpp — host CPU arithmetic for single-number entry points.
#include "host_math.h"

#include <algorithm>
#include <cstring>
#include <numeric>
#include <string>

#include "voss/voss.h"
#include "voss/voss_numbers.h"

namespace voss {

namespace {
inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }
uint64_t powmod(uint64_t a, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  a %= m;
  for (; e; e >>= 1) {
    if (e & 1) r = mulmod(r, a, m);
    a = mulmod(a, a, m);
  }
  return r;
}
}  // namespace

// miller_rabin.cpp:26-51: bases 2..37 decide every n < 3.3e24.
bool host_is_prime(uint64_t n) {
  if (n < 2) return false;
  if (n == 2 || n == 3) return true;
  if ((n & 1) == 0) return false;
  uint64_t d = n - 1;
  int r = 0;
  while ((d & 1) == 0) { d >>= 1; ++r; }
  static const uint64_t wit[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  for (uint64_t a : wit) {
    if (a >= n) continue;
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int k = 0; k < r - 1 && comp; ++k) {
      x = mulmod(x, x, n);
      if (x == n - 1) comp = false;
    }
    if (comp) return false;
  }
  return true;
}

// next_prev.cu:22-49: odd scan, at most 1e8 steps.
bool host_next_prime(uint64_t x, uint64_t* out) {
  if (x < 2) { *out = 2; return true; }
  uint64_t c = x + 1;
  if (c == 2) { *out = 2; return true; }
  if ((c & 1) == 0) ++c;
  for (uint64_t steps = 0; steps < 100000000ull; steps += 2, c += 2)
    if (host_is_prime(c)) { *out = c; return true; }
  return false;
}

// next_prev.cu:51-83 (caller rejects x <= 2).
bool host_prev_prime(uint64_t x, uint64_t* out) {
  if (x == 3) { *out = 2; return true; }
  uint64_t c = x - 1;
  if (c == 2) { *out = 2; return true; }
  if ((c & 1) == 0) --c;
  for (uint64_t steps = 0; steps < 100000000ull && c >= 3; steps += 2, c -= 2)
    if (host_is_prime(c)) { *out = c; return true; }
  if (c < 2) { *out = 2; return true; }
  return false;
}

namespace {
// Brent's variant of Pollard's rho (Brent 1980), returns a nontrivial factor of odd composite n.
uint64_t rho(uint64_t n) {
  for (uint64_t c = 1;; ++c) {
    uint64_t y = 2, x = 2, g = 1, q = 1, ys = 2;
    const uint64_t m = 128;
    uint64_t r = 1;
    auto f = [&](uint64_t v) { return (mulmod(v, v, n) + c) % n; };
    do {
      x = y;
      for (uint64_t i = 0; i < r; ++i) y = f(y);
      uint64_t k = 0;
      do {
        ys = y;
        for (uint64_t i = 0; i < std::min(m, r - k); ++i) {
          y = f(y);
          q = mulmod(q, x > y ? x - y : y - x, n);
        }
        g = std::gcd(q, n);
        k += m;
      } while (k < r && g == 1);
      r <<= 1;
    } while (g == 1);
    if (g == n) {
      do {
        ys = f(ys);
        g = std::gcd(x > ys ? x - ys : ys - x, n);
      } while (g == 1);
    }
    if (g != n) return g;
  }
}
void factor_rec(uint64_t n, std::vector<uint64_t>& out) {
  if (n == 1) return;
  if (host_is_prime(n)) { out.push_back(n); return; }
  const uint64_t d = rho(n);
  factor_rec(d, out);
  factor_rec(n / d, out);
}
}  // namespace

std::vector<uint64_t> host_factorize(uint64_t x) {
  std::vector<uint64_t> f;
  for (uint64_t p = 2; p < 100 && p * p <= x; ++p)
    while (x % p == 0) { f.push_back(p); x /= p; }
  if (x > 1) factor_rec(x, f);
  std::sort(f.begin(), f.end());
  return f;
}

std::vector<std::pair<uint64_t, uint32_t>> host_factor_exponents(uint64_t x) {
  std::vector<std::pair<uint64_t, uint32_t>> fe;
  for (uint64_t p : host_factorize(x)) {
    if (!fe.empty() && fe.back().first == p) fe.back().second++;
    else fe.push_back({p, 1});
  }
  return fe;
}

// Lucas-Lehmer: s_0 = 4, s_{k+1} = s_k^2 - 2 mod M_p; M_p prime iff s_{p-2} = 0 (p odd prime).
bool host_lucas_lehmer(uint32_t p) {
  if (p == 2) return true;
  const uint64_t M = (p == 64) ? ~0ull : ((1ull << p) - 1);
  uint64_t s = 4 % M;
  for (uint32_t k = 0; k + 2 < p; ++k) {
    const unsigned __int128 sq = (unsigned __int128)s * s + M - 2;
    s = (uint64_t)(sq % M);
  }
  return s == 0;
}

}  // namespace voss

// ---------------------------------------------------------------- voss_numbers.h
// Same domains and codes as library/src/numbers/*.cpp.
namespace voss { void numbers_set_err(const char* m); }  // capi.cpp: thread-local last error

using voss::host_factor_exponents;

extern "C" int voss_numbers_phi(uint64_t n, uint64_t* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n == 0) { voss::numbers_set_err("phi(0) is undefined; n must be >= 1"); return VOSS_ERR_INVALID_N; }
  uint64_t r = n;
  for (auto& pe : host_factor_exponents(n)) r = r / pe.first * (pe.first - 1);
  *out = r;
  return VOSS_OK;
}

extern "C" int voss_numbers_tau(uint64_t n, uint64_t* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n == 0) { voss::numbers_set_err("tau(0) is undefined; n must be >= 1"); return VOSS_ERR_INVALID_N; }
  uint64_t r = 1;
  for (auto& pe : host_factor_exponents(n)) r *= pe.second + 1;
  *out = r;
  return VOSS_OK;
}

extern "C" int voss_numbers_sigma(uint64_t n, uint64_t* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n == 0) { voss::numbers_set_err("sigma(0) is undefined; n must be >= 1"); return VOSS_ERR_INVALID_N; }
  unsigned __int128 r = 1;
  for (auto& pe : host_factor_exponents(n)) {
    unsigned __int128 term = 1, pw = 1;
    for (uint32_t i = 0; i < pe.second; ++i) { pw *= pe.first; term += pw; }
    r *= term;
    if (r > (unsigned __int128)UINT64_MAX) { voss::numbers_set_err("sigma overflows uint64"); return VOSS_ERR_OUT_OF_RANGE; }
  }
  *out = (uint64_t)r;
  return VOSS_OK;
}

extern "C" int voss_numbers_mu(uint64_t n, int* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n == 0) { voss::numbers_set_err("mu(0) is undefined; n must be >= 1"); return VOSS_ERR_INVALID_N; }
  int m = 1;
  for (auto& pe : host_factor_exponents(n)) {
    if (pe.second > 1) { *out = 0; return VOSS_OK; }
    m = -m;
  }
  *out = m;
  return VOSS_OK;
}

extern "C" int voss_numbers_gcd(uint64_t a, uint64_t b, uint64_t* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  *out = std::gcd(a, b);
  return VOSS_OK;
}

extern "C" int voss_numbers_lcm(uint64_t a, uint64_t b, uint64_t* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (a == 0 || b == 0) { *out = 0; return VOSS_OK; }
  const unsigned __int128 l = (unsigned __int128)(a / std::gcd(a, b)) * b;
  if (l > (unsigned __int128)UINT64_MAX) { voss::numbers_set_err("lcm overflow: (a/gcd)*b > UINT64_MAX"); return VOSS_ERR_OUT_OF_RANGE; }
  *out = (uint64_t)l;
  return VOSS_OK;
}

extern "C" int voss_numbers_fibonacci(uint64_t n, uint64_t* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n >= 94) { voss::numbers_set_err("fibonacci(n) overflows uint64 for n >= 94"); return VOSS_ERR_OUT_OF_RANGE; }
  uint64_t a = 0, b = 1;  // F(k), F(k+1) by fast doubling over the bits of n
  for (int bit = 63; bit >= 0; --bit) {
    const unsigned __int128 c = (unsigned __int128)a * (2 * (unsigned __int128)b - a);
    const unsigned __int128 d = (unsigned __int128)a * a + (unsigned __int128)b * b;
    a = (uint64_t)c;
    b = (uint64_t)d;
    if ((n >> bit) & 1) { const uint64_t t = a + b; a = b; b = t; }
  }
  *out = a;
  return VOSS_OK;
}

extern "C" int voss_numbers_factorial(uint64_t n, uint64_t* out) {
  if (!out) { voss::numbers_set_err("out pointer is null"); return VOSS_ERR_INVALID_ARG; }
  if (n >= 21) { voss::numbers_set_err("factorial(n) overflows uint64 for n >= 21"); return VOSS_ERR_OUT_OF_RANGE; }
  uint64_t r = 1;
  for (uint64_t i = 2; i <= n; ++i) r *= i;
  *out = r;
  return VOSS_OK;
}
