// The project's random stream, shared by the host GA driver and the device
// GA kernel so that both draw bit-identical values.
//
// State seeded by SplitMix64; xoshiro256** generator; 53-bit uniforms;
// Lemire's unbiased bounded integers; Box-Muller normals (one per call, the
// cosine branch).  The reference draws from rand's StdRng; the algorithms
// consuming the stream follow the reference step by step.
//
// The normal's logarithm and cosine are the stream's own functions, built
// from IEEE basic operations only (+, -, *, /, sqrt: correctly rounded on the
// host and on gfx950 alike, with FP contraction off; the square root is
// corrected to the rounded value), so every platform
// computes the same bits; library log and cos may differ in the last bit
// between the host's libm and the device's.  Both are within a few ulp of
// the true function.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

namespace gdock {
namespace stream {

__host__ __device__ inline uint64_t bits_of(double x) {
  uint64_t b;
  memcpy(&b, &x, 8);
  return b;
}

__host__ __device__ inline double from_bits(uint64_t b) {
  double x;
  memcpy(&x, &b, 8);
  return x;
}

// Natural logarithm of a positive normal x: x = m 2^e with m in
// [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s) with s = (m - 1) / (m + 1),
// |s| <= 0.1716, summed as 2s + 2s (z/3 + z^2/5 + ... + z^12/25), z = s^2
// (truncation below 1e-19 relative); e ln 2 with ln 2 split so that e ln2_hi
// is exact.
__host__ __device__ inline double ln(double x) {
  const uint64_t b = bits_of(x);
  int e = static_cast<int>((b >> 52) & 0x7ff) - 1023;
  double m = from_bits((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (m > 1.4142135623730951) {
    m = m * 0.5;  // exact
    e += 1;
  }
  const double f = m - 1.0;  // exact (m within a factor 2 of 1)
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 25.0;
  p = 1.0 / 23.0 + z * p;
  p = 1.0 / 21.0 + z * p;
  p = 1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = 1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = 1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = 1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = 1.0 / 3.0 + z * p;
  const double s2 = 2.0 * s;
  const double lm = s2 + s2 * (z * p);
  const double ln2_hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000: e ln2_hi exact for |e| < 2^11
  const double ln2_lo = 1.90821492927058770002e-10;
  const double de = static_cast<double>(e);
  return de * ln2_hi + (de * ln2_lo + lm);
}

// cos(2 pi u) for 0 <= u < 1: u = k/4 + r with k the nearest quarter
// (exact: r is on u's grid), |r| <= 1/8, x = 2 pi r (|x| <= pi/4), then the
// Taylor series of cos or sin to x^18 / x^19 (truncation below 1e-19) by the
// quadrant k mod 4.
__host__ __device__ inline double cos2pi(double u) {
  const double q4 = u * 4.0;  // exact
  const double k = static_cast<double>(static_cast<int64_t>(q4 + 0.5));
  const double r = u - k * 0.25;  // exact
  const double x = r * 6.283185307179586;
  const double z = x * x;
  double c = -1.0 / 6402373705728000.0;                      // -1/18!
  c = 1.0 / 20922789888000.0 + z * c;                        // 1/16!
  c = -1.0 / 87178291200.0 + z * c;                          // -1/14!
  c = 1.0 / 479001600.0 + z * c;                             // 1/12!
  c = -1.0 / 3628800.0 + z * c;                              // -1/10!
  c = 1.0 / 40320.0 + z * c;                                 // 1/8!
  c = -1.0 / 720.0 + z * c;                                  // -1/6!
  c = 1.0 / 24.0 + z * c;                                    // 1/4!
  c = -0.5 + z * c;
  const double cosx = 1.0 + z * c;
  double sn = -1.0 / 121645100408832000.0;                   // -1/19!
  sn = 1.0 / 355687428096000.0 + z * sn;                     // 1/17!
  sn = -1.0 / 1307674368000.0 + z * sn;                      // -1/15!
  sn = 1.0 / 6227020800.0 + z * sn;                          // 1/13!
  sn = -1.0 / 39916800.0 + z * sn;                           // -1/11!
  sn = 1.0 / 362880.0 + z * sn;                              // 1/9!
  sn = -1.0 / 5040.0 + z * sn;                               // -1/7!
  sn = 1.0 / 120.0 + z * sn;                                 // 1/5!
  sn = -1.0 / 6.0 + z * sn;                                  // -1/3!
  const double sinx = x + x * (z * sn);
  switch (static_cast<int>(k) & 3) {
    case 0: return cosx;
    case 1: return -sinx;
    case 2: return -cosx;
    default: return sinx;
  }
}

// Correctly rounded square root of a positive normal x: the platform's
// sqrt, then Tuckerman's test with exact FMA residuals moves it by one ulp
// if it is not the correctly rounded value.
__host__ __device__ inline double sqrt_cr(double x) {
  double y = sqrt(x);
  const double up = from_bits(bits_of(y) + 1), dn = from_bits(bits_of(y) - 1);
  if (fma(y, up, -x) < 0.0) {
    y = up;
  } else if (fma(dn, y, -x) >= 0.0) {
    y = dn;
  }
  return y;
}

struct Xoshiro256 {
  uint64_t s[4];

  __host__ __device__ explicit Xoshiro256(uint64_t seed = 0) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {  // SplitMix64
      x += 0x9e3779b97f4a7c15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      s[i] = z ^ (z >> 31);
    }
  }

  __host__ __device__ uint64_t next() {
    const uint64_t x = s[1] * 5;
    const uint64_t out = ((x << 7) | (x >> 57)) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
    return out;
  }

  __host__ __device__ double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0, 1)

  __host__ __device__ double range_closed(double lo, double hi) {  // [lo, hi]
    const double v = lo + (hi - lo) * (static_cast<double>(next() >> 11) / 9007199254740991.0);
    return v > hi ? hi : v;
  }

  // [0, n): Lemire's multiply-shift with rejection.
  __host__ __device__ uint64_t below(uint64_t n) {
    unsigned __int128 m = static_cast<unsigned __int128>(next()) * n;
    uint64_t low = static_cast<uint64_t>(m);
    if (low < n) {
      const uint64_t threshold = (0 - n) % n;
      while (low < threshold) {
        m = static_cast<unsigned __int128>(next()) * n;
        low = static_cast<uint64_t>(m);
      }
    }
    return static_cast<uint64_t>(m >> 64);
  }

  __host__ __device__ double normal(double mean, double sd) {  // Box-Muller, cosine branch
    const double u1 = 1.0 - unit(), u2 = unit();
    return mean + sd * (sqrt_cr(-2.0 * ln(u1)) * cos2pi(u2));
  }
};

// Jump-ahead.  xoshiro256's state transition T is linear over GF(2) (shifts,
// a rotation and XORs of the four words), so the state after j steps is
// q(T) s with q = x^j mod P, P the characteristic polynomial of T (degree
// 256): the XOR of the states T^i s over the set bits i of q.  P comes from
// Berlekamp-Massey on the sequence of one state bit, and is checked by
// P(T) s = 0; jump_polys() tabulates q for j = kJumpStride k.
constexpr int kJumpStride = 16;
constexpr int kJumpEntries = 2048;  // jumps of up to 16 x 2047 steps

// XOR of base[i] (the states T^i s, i < 256) over the set bits of q.
__host__ __device__ inline void jump_combine(const uint64_t (*base)[4], const uint64_t* q, uint64_t (&out)[4]) {
  out[0] = out[1] = out[2] = out[3] = 0;
  for (int w = 0; w < 4; ++w) {
    uint64_t bits = q[w];
    while (bits) {
      const int i = 64 * w + __builtin_ctzll(bits);
      bits &= bits - 1;
      out[0] ^= base[i][0];
      out[1] ^= base[i][1];
      out[2] ^= base[i][2];
      out[3] ^= base[i][3];
    }
  }
}

// P without its x^256 term (bit i: coefficient of x^i); false unless the
// sequence's linear complexity is 256 and P(T) annihilates the state.
inline bool char_poly(uint64_t (&p)[4]) {
  constexpr int N = 512, L_MAX = 256;
  Xoshiro256 r(0x5eed5eed5eedull);
  const Xoshiro256 s0 = r;
  uint8_t seq[N];
  for (int i = 0; i < N; ++i) {  // bit 0 of word 0 along the orbit
    seq[i] = static_cast<uint8_t>(r.s[0] & 1u);
    r.next();
  }
  uint8_t c[N + 1] = {1}, b[N + 1] = {1}, t[N + 1];  // Berlekamp-Massey over GF(2)
  int L = 0, m = 1;
  for (int n = 0; n < N; ++n) {
    uint8_t d = seq[n];
    for (int i = 1; i <= L; ++i) d ^= static_cast<uint8_t>(c[i] & seq[n - i]);
    if (!d) {
      ++m;
      continue;
    }
    memcpy(t, c, sizeof t);
    for (int i = 0; i + m <= N; ++i) c[i + m] ^= b[i];
    if (2 * L <= n) {
      L = n + 1 - L;
      memcpy(b, t, sizeof b);
      m = 1;
    } else {
      ++m;
    }
  }
  if (L != L_MAX) return false;
  p[0] = p[1] = p[2] = p[3] = 0;
  for (int k = 0; k < L_MAX; ++k)  // P_k = c_(L-k)
    if (c[L_MAX - k]) p[k / 64] |= 1ull << (k % 64);
  Xoshiro256 w = s0;  // P(T) s = sum of P_k T^k s, with P_256 = 1
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int k = 0; k <= L_MAX; ++k) {
    if (k == L_MAX || ((p[k / 64] >> (k % 64)) & 1u))
      for (int q = 0; q < 4; ++q) acc[q] ^= w.s[q];
    w.next();
  }
  return (acc[0] | acc[1] | acc[2] | acc[3]) == 0;
}

// table[4 k .. 4 k + 3] = x^(kJumpStride k) mod P, for k < kJumpEntries.
inline bool jump_polys(uint64_t* table) {
  uint64_t p[4];
  if (!char_poly(p)) return false;
  uint64_t q[4] = {1, 0, 0, 0};
  for (int k = 0; k < kJumpEntries; ++k) {
    memcpy(table + 4 * k, q, sizeof q);
    for (int s = 0; s < kJumpStride; ++s) {  // q <- x q mod P
      const bool top = (q[3] >> 63) != 0;
      q[3] = (q[3] << 1) | (q[2] >> 63);
      q[2] = (q[2] << 1) | (q[1] >> 63);
      q[1] = (q[1] << 1) | (q[0] >> 63);
      q[0] <<= 1;
      if (top)
        for (int w = 0; w < 4; ++w) q[w] ^= p[w];
    }
  }
  return true;
}

}  // namespace stream
}  // namespace gdock
