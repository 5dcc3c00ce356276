// gnx_rng.cpp -- the reference's host random stream (include/gnx_rng.h).
//
// dSFMT (Saito & Matsumoto, "SIMD-oriented Fast Mersenne Twister for double
// precision", 2009) with the MEXP = 132049 parameter set the reference
// selects (myrngparam.h:4, dSFMT-params132049.h), written from the published
// algorithm in plain 64-bit integer arithmetic: the state is N + 1 = 1270
// 128-bit words held as pairs of uint64 (the last one the "lung"), refilled
// N words at a time by the recursion
//     lung' = (a << SL1) ^ swap32(lung) ^ b,
//     r     = (lung' >> SR) ^ (lung' & MSK) ^ a,
// with b the word POS1 ahead; each payload word is an IEEE double in [1, 2)
// and open_open maps it to (0, 1) by setting the lowest mantissa bit and
// subtracting 1.  Seeding fills the state's 32-bit words with the
// 1812433253 recurrence, forces the payload into [1, 2) and certifies the
// period on the lung word.  The draw helpers follow random.c line by line
// (the rejection loops, v = 1 - 2u or -1 + 2u, and the scaling), so that the
// stream of values is the reference's.
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#include <new>

#include "../../include/gnx_rng.h"

namespace {

// dSFMT-params132049.h
constexpr int kN = 1269;  // 128-bit payload words: (MEXP - 128) / 104 + 1
constexpr int kPos1 = 371;
constexpr int kSL1 = 23;
constexpr int kSR = 12;
constexpr uint64_t kMsk[2] = {0x000fb9f4eff4bf77ull, 0x000fffffbfefff37ull};
constexpr uint64_t kFix[2] = {0x4ce24c0e4e234f3bull, 0x62612409b5665c2dull};
constexpr uint64_t kPcv[2] = {0x181232889145d000ull, 0x0000000000000001ull};
constexpr uint64_t kLowMask = 0x000FFFFFFFFFFFFFull;
constexpr uint64_t kOne = 0x3FF0000000000000ull;  // the exponent of [1, 2)

struct Stream {
  uint64_t w[2 * (kN + 1)];  // payload words 0..N-1, lung word N
  int next;                  // next payload 64-bit half to hand out

  void seed(uint32_t s) {
    // the state's 32-bit words, low half of each 64-bit lane first
    uint32_t prev = s;
    for (int i = 0; i < 4 * (kN + 1); i++) {
      const uint32_t v = i == 0 ? s : 1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i;
      prev = v;
      uint64_t& lane = w[i >> 1];
      lane = (i & 1) ? (lane | ((uint64_t)v << 32)) : (uint64_t)v;
    }
    for (int i = 0; i < 2 * kN; i++) w[i] = (w[i] & kLowMask) | kOne;
    // period certification: the parity of (lung ^ fix) & pcv must be odd
    uint64_t inner = ((w[2 * kN] ^ kFix[0]) & kPcv[0]) ^ ((w[2 * kN + 1] ^ kFix[1]) & kPcv[1]);
    for (int sh = 32; sh > 0; sh >>= 1) inner ^= inner >> sh;
    if (!(inner & 1)) w[2 * kN + 1] ^= 1;  // kPcv[1] is 1: flip its bit
    next = 2 * kN;
  }

  void refill() {
    uint64_t l0 = w[2 * kN], l1 = w[2 * kN + 1];
    for (int i = 0; i < kN; i++) {
      const int j = i + kPos1 < kN ? i + kPos1 : i + kPos1 - kN;
      const uint64_t a0 = w[2 * i], a1 = w[2 * i + 1];
      const uint64_t n0 = (a0 << kSL1) ^ (l1 >> 32) ^ (l1 << 32) ^ w[2 * j];
      const uint64_t n1 = (a1 << kSL1) ^ (l0 >> 32) ^ (l0 << 32) ^ w[2 * j + 1];
      l0 = n0;
      l1 = n1;
      w[2 * i] = (n0 >> kSR) ^ (n0 & kMsk[0]) ^ a0;
      w[2 * i + 1] = (n1 >> kSR) ^ (n1 & kMsk[1]) ^ a1;
    }
    w[2 * kN] = l0;
    w[2 * kN + 1] = l1;
    next = 0;
  }

  // dsfmt_genrand_open_open
  double open_open() {
    if (next >= 2 * kN) refill();
    double d;
    const uint64_t bits = w[next++] | 1;
    memcpy(&d, &bits, sizeof d);
    return d - 1.0;
  }

  // random.c:88-102
  double gauss1() {
    double v1, v2, s;
    do {
      v1 = 1.0 - 2.0 * open_open();
      v2 = 1.0 - 2.0 * open_open();
      s = v1 * v1 + v2 * v2;
    } while (s >= 1);
    return v1 * sqrt(-2 * log(s) / s);
  }

  // random.c:129-143
  void gauss2mv(double mean, double variance, double* r1, double* r2) {
    double v1, v2, s;
    do {
      v1 = -1.0 + 2.0 * open_open();
      v2 = -1.0 + 2.0 * open_open();
      s = v1 * v1 + v2 * v2;
    } while (s >= 1);
    *r1 = mean + sqrt(variance) * v1 * sqrt(-2 * log(s) / s);
    *r2 = mean + sqrt(variance) * v2 * sqrt(-2 * log(s) / s);
  }

  double z2() { return open_open() <= 0.5 ? 1.0 : -1.0; }
};

}  // namespace

struct gnx_rng {
  Stream s;
};

extern "C" {

gnx_status gnx_rng_create(unsigned int seed, gnx_rng** out) {
  if (!out) return GNX_ERR_ARG;
  gnx_rng* r = new (std::nothrow) gnx_rng;
  if (!r) return GNX_ERR_SIZE;
  // Initrand (random.c:55-62): 0 selects a clock-based seed
  r->s.seed(seed ? seed : (unsigned int)(((unsigned long)time(nullptr) + 10) % 0xFFFFFFFFul));
  *out = r;
  return GNX_OK;
}

gnx_status gnx_rng_destroy(gnx_rng* r) {
  delete r;
  return GNX_OK;
}

double gnx_rng_casuale(gnx_rng* r) { return r->s.open_open(); }

double gnx_rng_gauss1(gnx_rng* r) { return r->s.gauss1(); }

void gnx_rng_gauss2mv(gnx_rng* r, double mean, double variance, double* r1, double* r2) {
  r->s.gauss2mv(mean, variance, r1, r2);
}

gnx_status gnx_rng_fill(gnx_rng* r, gnx_draw kind, long n, double var, double* out) {
  if (!r || !out || n < 0) return GNX_ERR_ARG;
  Stream& s = r->s;
  switch (kind) {
    case GNX_DRAW_CASUALE:
      for (long i = 0; i < n; i++) out[i] = s.open_open();
      return GNX_OK;
    case GNX_DRAW_GAUSS1:
      for (long i = 0; i < n; i++) out[i] = s.gauss1();
      return GNX_OK;
    case GNX_DRAW_SIGMA: {
      if (n % 2) return GNX_ERR_ARG;
      const long half = n / 2;
      for (long x = 0; x < half; x++) s.gauss2mv(0.0, var, &out[x], &out[x + half]);
      return GNX_OK;
    }
    case GNX_DRAW_Z2_COMPLEX:
      for (long i = 0; i < 2 * n; i++) out[i] = s.z2();
      return GNX_OK;
    case GNX_DRAW_Z2_REAL:
      for (long i = 0; i < n; i++) out[i] = s.z2();
      return GNX_OK;
    case GNX_DRAW_GAUSS2_PAIRS:
      if (n % 2) return GNX_ERR_ARG;
      for (long i = 0; i < n; i += 2) s.gauss2mv(0.0, var, &out[i], &out[i + 1]);
      return GNX_OK;
  }
  return GNX_ERR_ARG;
}

}  // extern "C"
