// Arithmetic primitives shared by the engine's host code and its gfx950 kernels.
//
//  * Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw, "Parallel random
//    numbers: as easy as 1, 2, 3", SC'11).  The reference draws from cuRAND's
//    curandStatePhilox4_32_10_t (PAisingSSC.cu:33, PAisingMSC.cu:32): curand_init(seed,
//    subsequence, 0) puts the seed in the key and the subsequence in counter words 2-3, and
//    curand() returns the four output words of counter block k as stream words 4k..4k+3.
//    Word p of a stream is therefore a pure function of (seed, subsequence, p).
//  * The uniform double of curand_uniform_double: one 32-bit word v per double, mapped to
//    (v+1)/2^32 (matches the CUDA >= 10.1 reference outputs) or (v+1/2)/2^32 (CUDA <= 10.0).
//  * Own FP64 exp/log (Cody-Waite reduction + Taylor/atanh series) so that the device and
//    the host compute lnQ and the copy counts bit-identically.  The operation sequence is
//    fixed; compile with -ffp-contract=off and keep every fma explicit.
#ifndef PA_PRIMS_H
#define PA_PRIMS_H

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PA_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define PA_HD inline
#endif

namespace pa {

constexpr uint32_t kLcgA = 1664525u;     // PAisingMSC.cu:53
constexpr uint32_t kLcgC = 1013904223u;  // PAisingMSC.cu:54
constexpr int kStreamsPerUnit = 128;     // EQthreads, PAisingSSC.cu:62: streams per replica/group
constexpr int kRedBlock = 1024;          // Nthreads, PAisingSSC.cu:63: reduction block size

enum UniformMap : int { kUniformOpenClosed = 0, kUniformMidpoint = 1 };

struct U4 {
  uint32_t x, y, z, w;
};

PA_HD uint32_t mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * b) >> 32);
}

// a ^ b ^ c; one v_bitop3_b32 (truth table 0x96) on gfx950.
PA_HD uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// Ten Philox rounds on counter c with key (k0, k1).
PA_HD U4 philox10(U4 c, uint32_t k0, uint32_t k1) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    U4 n;
    n.x = xor3((uint32_t)(p1 >> 32), c.y, k0);
    n.y = (uint32_t)p1;
    n.z = xor3((uint32_t)(p0 >> 32), c.w, k1);
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Block `blk` (words 4*blk .. 4*blk+3) of subsequence `sub` under `seed`.
PA_HD U4 stream_block(uint64_t seed, uint64_t sub, uint64_t blk) {
  U4 c{(uint32_t)blk, (uint32_t)(blk >> 32), (uint32_t)sub, (uint32_t)(sub >> 32)};
  return philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Block `blk` of the four subsequences sub .. sub+3.  When they share their high 32 bits
// (all but three in 2^32 cases) the four counters differ only in their third word, so the
// first round's product of the block index and xor of the high word, and the second round's
// product of the third word, are common to the four and evaluated once; otherwise each
// block is evaluated on its own.  The results equal four stream_block calls.
PA_HD void stream_block_quad(uint64_t seed, uint64_t sub, uint64_t blk, U4 (&o)[4]) {
  const uint32_t lo = (uint32_t)sub, hi = (uint32_t)(sub >> 32);
  if (lo <= 0xFFFFFFFCu) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int j = 0; j < 4; ++j)
      o[j] = philox10(U4{(uint32_t)blk, (uint32_t)(blk >> 32), lo + (uint32_t)j, hi},
                      (uint32_t)seed, (uint32_t)(seed >> 32));
  } else {
    for (int j = 0; j < 4; ++j) o[j] = stream_block(seed, sub + (uint64_t)j, blk);
  }
}

PA_HD uint32_t u4_word(const U4& v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

PA_HD double uniform_double(uint32_t v, int umap) {
  const double inv32 = 2.3283064365386963e-10;  // 2^-32
  return umap == kUniformOpenClosed ? inv32 + (double)v * inv32 : (double)v * inv32 + inv32 * 0.5;
}

PA_HD double bits_to_double(uint64_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __longlong_as_double((long long)u);
#else
  union { uint64_t u; double d; } c;
  c.u = u;
  return c.d;
#endif
}

PA_HD uint64_t double_to_bits(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (uint64_t)__double_as_longlong(d);
#else
  union { uint64_t u; double d; } c;
  c.d = d;
  return c.u;
#endif
}

PA_HD double fma_(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_fma(a, b, c);
#else
  return std::fma(a, b, c);
#endif
}

PA_HD double rint_(double a) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_rint(a);
#else
  return std::nearbyint(a);
#endif
}

// exp(x): x = k ln2 + r, |r| <= ln2/2 (Cody & Waite), exp(r) = 1 + (r + r^2 P(r)) with P
// the degree-11 Taylor tail 1/2! + r/3! + ... + r^11/13!, then scaled by 2^k.
PA_HD double exp_fp64(double x) {
  if (x != x) return x;
  if (x > 709.79) return bits_to_double(0x7FF0000000000000ull);
  if (x < -745.2) return 0.0;
  const double k = rint_(x * 1.4426950408889634);
  double r = fma_(-k, 0.6931471805599453, x);
  r = fma_(-k, 2.3190468138462996e-17, r);
  double p = 1.6059043836821613e-10;
  p = fma_(p, r, 2.08767569878681e-09);
  p = fma_(p, r, 2.505210838544172e-08);
  p = fma_(p, r, 2.755731922398589e-07);
  p = fma_(p, r, 2.7557319223985893e-06);
  p = fma_(p, r, 2.48015873015873e-05);
  p = fma_(p, r, 0.0001984126984126984);
  p = fma_(p, r, 0.001388888888888889);
  p = fma_(p, r, 0.008333333333333333);
  p = fma_(p, r, 0.041666666666666664);
  p = fma_(p, r, 0.16666666666666666);
  p = fma_(p, r, 0.5);
  const double rr = r * r;
  const double t = fma_(rr, p, r);
  double y = 1.0 + t;
  int ki = (int)k;
  if (ki > 1023) { y *= 2.0; ki -= 1; }
  if (ki < -1022) {
    y *= bits_to_double((uint64_t)(ki + 600 + 1023) << 52);
    return y * bits_to_double((uint64_t)(1023 - 600) << 52);
  }
  return y * bits_to_double((uint64_t)(ki + 1023) << 52);
}

// log(x): x = 2^e m, m in (sqrt(1/2), sqrt(2)], f = m - 1, s = f/(2+f),
// log(1+f) = f - (f^2/2 - s (f^2/2 + R)), R = sum_{k=1..11} 2 s^{2k}/(2k+1).
PA_HD double log_fp64(double x) {
  if (x != x) return x;
  if (x < 0.0) return bits_to_double(0x7FF8000000000000ull);
  if (x == 0.0) return bits_to_double(0xFFF0000000000000ull);
  if (x == bits_to_double(0x7FF0000000000000ull)) return x;
  int e = 0;
  if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; e = -54; }
  const uint64_t b = double_to_bits(x);
  e += (int)(b >> 52) - 1023;
  double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 0.08695652173913043;
  R = R * z + 0.09523809523809523;
  R = R * z + 0.10526315789473684;
  R = R * z + 0.11764705882352941;
  R = R * z + 0.13333333333333333;
  R = R * z + 0.15384615384615385;
  R = R * z + 0.18181818181818182;
  R = R * z + 0.2222222222222222;
  R = R * z + 0.2857142857142857;
  R = R * z + 0.4;
  R = R * z + 0.6666666666666666;
  R = R * z;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)e;
  return dk * 0.6931471803691238 - ((hfsq - (s * (hfsq + R) + dk * 1.9082149292705877e-10)) - f);
}

// splitmix64 finaliser (Steele, Lea, Flood 2014), used by the order-independent
// population hashes sum_i mix64(mix64(i) ^ chunk_i).
PA_HD uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace pa

#endif
