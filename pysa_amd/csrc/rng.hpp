// Counter-based random numbers and the deterministic logarithm used by every
// kernel of the engine.
//
// * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As
//   Easy as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies, with the
//   key bumped by the Weyl constants between rounds.
// * Uniforms are u = (x + 0.5) * 2^-32 for a 32-bit word x, so u lies strictly
//   inside (0, 1) and log(u) is always finite.
// * det_log(x) is a logarithm written only with IEEE-754 +, -, *, / and fma in
//   a fixed order.  Compiled without fast-math and with -ffp-contract=off it
//   gives the same bits on the CPU and on the GPU, which is what makes the
//   engine's trajectories reproducible bit for bit on both sides.
//   Method: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//   ln m = 2*atanh(s) = 2s(1 + s^2/3 + s^4/5 + ...), ln x = e*ln2 + ln m.
// * Feistel permutation: a 4-round balanced Feistel network on the smallest
//   even number of bits covering n, with cycle walking, gives a keyed
//   bijection of [0, n) that any lane can evaluate independently.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PYSA_HD __host__ __device__ __forceinline__
#else
#define PYSA_HD static inline
#endif

namespace pysa {

// Stream purposes, kept in the top byte of counter word 2.
enum : uint32_t {
  kStreamSweep = 0u,   // one uniform per (chain, sweep, spin position)
  kStreamPT = 1u,      // one uniform per (read, sweep, replica pair)
  kStreamPerm = 2u,    // Feistel keys per (chain, sweep)
  kStreamShared = 3u,  // Feistel keys per sweep, shared by all chains
  kStreamBlock = 4u,   // block-random order: block permutation keys per sweep
  kStreamMask = 5u,    // block-random order: in-block XOR masks per (sweep, block)
};

// Spins per block of the block-random visit order (and of the MFMA form).
constexpr int kOrderBlock = 32;

struct u32x4 {
  uint32_t x, y, z, w;
};

PYSA_HD void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  *lo = (uint32_t)p;
}

PYSA_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(M0, c.x, &hi0, &lo0);
    mulhilo32(M1, c.z, &hi1, &lo1);
    u32x4 n;
    n.x = hi1 ^ c.y ^ k0;
    n.y = lo1;
    n.z = hi0 ^ c.w ^ k1;
    n.w = lo0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

PYSA_HD uint32_t pick_word(u32x4 v, uint32_t i) {
  return (i & 2u) ? ((i & 1u) ? v.w : v.z) : ((i & 1u) ? v.y : v.x);
}

// Uniform in (0,1) from a 32-bit word; exact in double.
PYSA_HD double u01_from_u32(uint32_t x) {
  return ((double)x + 0.5) * 2.3283064365386962890625e-10;  // 2^-32
}

PYSA_HD double as_double(uint64_t b) {
  union { uint64_t u; double d; } v;
  v.u = b;
  return v.d;
}
PYSA_HD uint64_t as_u64(double d) {
  union { uint64_t u; double d; } v;
  v.d = d;
  return v.u;
}

PYSA_HD double dfma(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return __builtin_fma(a, b, c);
#endif
}

// Deterministic natural logarithm of a positive, normal, finite double.
PYSA_HD double det_log(double x) {
  // ln 2 split into a 29-bit head (so e*LN2_HI is exact for |e| < 2^24) and
  // the double nearest to the remainder.
  const double LN2_HI = 0.6931471787393093;     // 0x3FE62E42FE000000
  const double LN2_LO = 1.8206359985041462e-09;
  uint64_t b = as_u64(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = as_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) {  // fold into [sqrt(1/2), sqrt(2))
    m = m * 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double z = s * s;
  // p(z) = sum_{k>=1} z^(k-1) / (2k+1), k = 1..12, Horner from the top.
  double p = 1.0 / 25.0;
  p = dfma(p, z, 1.0 / 23.0);
  p = dfma(p, z, 1.0 / 21.0);
  p = dfma(p, z, 1.0 / 19.0);
  p = dfma(p, z, 1.0 / 17.0);
  p = dfma(p, z, 1.0 / 15.0);
  p = dfma(p, z, 1.0 / 13.0);
  p = dfma(p, z, 1.0 / 11.0);
  p = dfma(p, z, 1.0 / 9.0);
  p = dfma(p, z, 1.0 / 7.0);
  p = dfma(p, z, 1.0 / 5.0);
  p = dfma(p, z, 1.0 / 3.0);
  double s2 = s + s;
  double lnm = dfma(s2 * z, p, s2);  // 2s + 2s*z*p
  double de = (double)e;
  return dfma(de, LN2_HI, dfma(de, LN2_LO, lnm));
}

// log of a Philox word, as the engine's log-uniform.
PYSA_HD double log_u32(uint32_t x) { return det_log(u01_from_u32(x)); }

// MurmurHash3 32-bit finaliser (public domain), used as the Feistel round
// function.
PYSA_HD uint32_t fmix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

struct FeistelKey {
  uint32_t k[4];
  uint32_t half_bits;  // h: the network permutes [0, 2^(2h))
  uint32_t n;
};

PYSA_HD FeistelKey feistel_make(uint32_t n, u32x4 keys) {
  FeistelKey f;
  f.k[0] = keys.x;
  f.k[1] = keys.y;
  f.k[2] = keys.z;
  f.k[3] = keys.w;
  f.n = n;
  uint32_t bits = 0;
  while (bits < 32 && (1u << bits) < n) ++bits;
  f.half_bits = (bits + 1) / 2;
  return f;
}

PYSA_HD uint32_t feistel_round_trip(const FeistelKey& f, uint32_t x) {
  const uint32_t h = f.half_bits;
  const uint32_t mask = (h >= 32) ? 0xFFFFFFFFu : ((1u << h) - 1u);
  uint32_t L = x >> h, R = x & mask;
  for (int i = 0; i < 4; ++i) {
    uint32_t F = fmix32(R ^ f.k[i]) & mask;
    uint32_t nl = R;
    R = L ^ F;
    L = nl;
  }
  return (L << h) | R;
}

// The i-th visited position of the keyed permutation of [0, n).
PYSA_HD uint32_t feistel_perm(const FeistelKey& f, uint32_t i) {
  if (f.n <= 1) return 0;
  uint32_t y = feistel_round_trip(f, i);
  while (y >= f.n) y = feistel_round_trip(f, y);
  return y;
}

// Counter layouts (key = seed words); chain < 2^40 and sweep < 2^48, with
// chain_lo = chain bits 0-23, chain_hi = bits 24-39, sweep_lo / sweep_hi =
// bits 0-31 / 32-47:
//   sweep uniforms:  (pos >> 2, sweep_lo, chain_lo | purpose<<24, chain_hi | sweep_hi<<16)
//   PT uniforms:     (pair >> 2, sweep_lo, purpose<<24 | sweep_hi, read)
//   Feistel keys:    (0, sweep_lo, chain_lo | purpose<<24, chain_hi | sweep_hi<<16)
PYSA_HD uint32_t chain_sweep_hi(uint64_t chain, uint64_t sweep) {
  return (uint32_t)((chain >> 24) & 0xFFFFu) | ((uint32_t)((sweep >> 32) & 0xFFFFu) << 16);
}
PYSA_HD u32x4 sweep_block(uint64_t chain, uint64_t sweep, uint32_t pos, uint32_t k0, uint32_t k1) {
  u32x4 c;
  c.x = pos >> 2;
  c.y = (uint32_t)sweep;
  c.z = ((uint32_t)chain & 0x00FFFFFFu) | (kStreamSweep << 24);
  c.w = chain_sweep_hi(chain, sweep);
  return philox4x32_10(c, k0, k1);
}

PYSA_HD u32x4 pt_block(uint64_t read, uint64_t sweep, uint32_t pair, uint32_t k0, uint32_t k1) {
  u32x4 c;
  c.x = pair >> 2;
  c.y = (uint32_t)sweep;
  c.z = (kStreamPT << 24) | (uint32_t)((sweep >> 32) & 0xFFFFFFu);
  c.w = (uint32_t)read;
  return philox4x32_10(c, k0, k1);
}

PYSA_HD u32x4 perm_keys(uint64_t chain, uint64_t sweep, uint32_t purpose, uint32_t k0, uint32_t k1) {
  u32x4 c;
  c.x = 0;
  c.y = (uint32_t)sweep;
  c.z = ((uint32_t)chain & 0x00FFFFFFu) | (purpose << 24);
  c.w = chain_sweep_hi(chain, sweep);
  return philox4x32_10(c, k0, k1);
}

// Block-random visit order of one sweep, shared by every chain of a run:
// the nblk = ceil(n / 32) blocks of 32 consecutive positions are visited in
// the order of a keyed Feistel permutation of [0, nblk), and inside block b
// the 32 slots k = 0..31 visit position 32 b + (k ^ m_b) for a 5-bit mask
// m_b (positions >= n are skipped).  The order does not depend on the state,
// so each single-spin Metropolis update keeps the Boltzmann distribution
// stationary, as with any random scan.
//   block keys:  perm_keys(0, sweep, kStreamBlock)
//   masks:       word b & 3 of Philox((b >> 2, sweep_lo, kStreamMask << 24, sweep_hi)), low 5 bits
PYSA_HD FeistelKey block_order_key(uint32_t nblk, uint64_t sweep, uint32_t k0, uint32_t k1) {
  return feistel_make(nblk, perm_keys(0ull, sweep, kStreamBlock, k0, k1));
}
PYSA_HD uint32_t block_mask(uint64_t sweep, uint32_t b, uint32_t k0, uint32_t k1) {
  u32x4 c;
  c.x = b >> 2;
  c.y = (uint32_t)sweep;
  c.z = kStreamMask << 24;
  c.w = (uint32_t)((sweep >> 32) & 0xFFFFu);
  return pick_word(philox4x32_10(c, k0, k1), b & 3u) & 31u;
}
// Position visited at step i (0 <= i < 32 * nblk) of the sweep, or -1 for a
// padding slot of the last block.
PYSA_HD int block_random_pos(const FeistelKey& f, uint64_t sweep, uint32_t i, int n, uint32_t k0,
                             uint32_t k1) {
  const uint32_t b = feistel_perm(f, i >> 5);
  const int pos = (int)(32u * b + ((i & 31u) ^ block_mask(sweep, b, k0, k1)));
  return pos < n ? pos : -1;
}

}  // namespace pysa
