// Multiply-with-carry multiplier search.
//
// A lag-1 MWC generator x' = (a*x + c) mod 2^32, c' = (a*x + c) >> 32 has the
// maximal period (a*2^32 - 2)/2 when p = a*2^32 - 1 is a safe prime, i.e. both
// p and (p-1)/2 are prime.  The ice directories ship these multipliers in
// rnd.txt (read by ini.cxx:767-799, one stream per multiplier); the shipped
// table is the descending sequence of every a < 2^32 with that property.  This
// file regenerates that sequence (and continues it below the shipped table when
// a launch needs more concurrent streams than the file provides).
//
// Method: a segmented sieve removes a with p or (p-1)/2 divisible by a small
// odd prime, then a deterministic Miller-Rabin test (bases 2..37, exact for
// all n < 3.3e24) settles the survivors.
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "xppc_c.h"

namespace {

typedef unsigned __int128 u128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod(r, b, m);
    b = mulmod(b, b, m);
    e >>= 1;
  }
  return r;
}

bool is_prime_u64(uint64_t n) {
  static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  if (n < 2) return false;
  for (uint64_t p : small) {
    if (n % p == 0) return n == p;
  }
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) d >>= 1, s++;
  for (uint64_t a : small) {
    uint64_t x = powmod(a, d, n);
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int r = 1; r < s; r++) {
      x = mulmod(x, x, n);
      if (x == n - 1) { comp = false; break; }
    }
    if (comp) return false;
  }
  return true;
}

// modular inverse of v modulo odd prime q (v not divisible by q)
uint32_t inv_mod(uint64_t v, uint32_t q) { return (uint32_t)powmod(v % q, q - 2, q); }

}  // namespace

namespace {

struct Sieve {
  std::vector<uint32_t> primes, r1, r2;
  Sieve() {
    const uint32_t QMAX = 1 << 16;
    std::vector<char> comp(QMAX + 1, 0);
    for (uint32_t i = 2; i <= QMAX; i++) {
      if (comp[i]) continue;
      if (i > 2) primes.push_back(i);
      for (uint64_t j = (uint64_t)i * i; j <= QMAX; j += i) comp[j] = 1;
    }
    // residues of a to strike: a*2^32 == 1 (q | p) and a*2^31 == 1 (q | (p-1)/2)
    r1.resize(primes.size());
    r2.resize(primes.size());
    for (size_t k = 0; k < primes.size(); k++) {
      uint32_t q = primes[k];
      r1[k] = inv_mod((1ULL << 32) % q, q);
      r2[k] = inv_mod((1ULL << 31) % q, q);
    }
  }

  // all multipliers in [lo, hi], descending
  void block(int64_t lo, int64_t hi, std::vector<uint32_t>& res) const {
    uint32_t len = (uint32_t)(hi - lo + 1);
    std::vector<char> bad(len, 0);
    for (size_t k = 0; k < primes.size(); k++) {
      uint32_t q = primes[k];
      uint32_t base = (uint32_t)(lo % q);
      for (uint32_t r : {r1[k], r2[k]}) {
        for (uint32_t i = (r + q - base) % q; i < len; i += q) bad[i] = 1;
      }
    }
    for (int64_t a = hi; a >= lo; a--) {
      if (bad[a - lo]) continue;
      uint64_t p = ((uint64_t)a << 32) - 1;
      if (is_prime_u64((p - 1) / 2) && is_prime_u64(p)) res.push_back((uint32_t)a);
    }
  }
};

}  // namespace

extern "C" int ppc_mwc_multipliers(uint32_t below, uint32_t n, uint32_t* out) {
  if (n == 0) return 0;
  if (out == nullptr) return -1;
  static const Sieve sv;
  const int64_t BLK = 1 << 20;
  unsigned nth = std::thread::hardware_concurrency();
  if (nth < 1) nth = 1;
  if (nth > 16) nth = 16;
  uint32_t found = 0;
  int64_t hi = (int64_t)below - 1;  // candidates a <= hi, descending
  while (found < n && hi >= 2) {
    // one wave of nth consecutive blocks, searched in parallel, merged in order
    std::vector<std::vector<uint32_t>> res(nth);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; t++) {
      int64_t bhi = hi - (int64_t)t * BLK, blo = bhi - BLK + 1;
      if (blo < 2) blo = 2;
      if (bhi < 2) break;
      pool.emplace_back([&, t, blo, bhi] { sv.block(blo, bhi, res[t]); });
    }
    for (auto& th : pool) th.join();
    for (unsigned t = 0; t < pool.size(); t++)
      for (uint32_t a : res[t])
        if (found < n) out[found++] = a;
    hi -= (int64_t)nth * BLK;
  }
  return (int)found;
}
