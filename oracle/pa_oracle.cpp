/*
 * CPU oracle: host C++/OpenMP restatement of the reference PA programs.
 *
 * TEST INFRASTRUCTURE ONLY (checker for the HIP engine; see pa_oracle.h).
 *
 * Citation keys: SSC = PAisingSSC.cu, MSC = PAisingMSC.cu (LevBarash/PAising).
 *
 * Structure follows the reference exactly where results depend on it:
 *  - virtual thread blocks of EQthreads=128 / Nthreads=1024 threads (SSC:62-63),
 *    because they fix which Philox subsequence drives which site;
 *  - the warp-32 shuffle-down / block reduction tree (SSC:107-124), because it fixes
 *    the FP64 summation order of the partition-function ratio q;
 *  - the host beta loop, lnQ and thermodynamics arithmetic (SSC:508-664).
 *
 * Pinning (how each routine is checked against something other than itself):
 *  - ora_philox: Salmon et al. SC'11 known-answer vectors (tests/test_oracle_rng.py)
 *    and rocRAND's MIT header compiled host-side in a test.
 *  - stream layout, init, energy, Q: reference fixture rung lnQ_1 (MSC) and lnQ_1 (SSC).
 *  - uniform-double mapping: fixture rungs R_1 and E_1 (SSC and MSC).
 *  - sweeps (SSC, MSC), resampling order, averages, writer: all 200 fixture rows.
 *  - energy/magnetisation: exhaustive brute force on tiny lattices.
 *  - ora_exp / ora_log: against libm (<= 1 ulp) and special cases.
 *  - MSC with 64 replicas per word: init word order is unspecified in the reference
 *    (MSC:183-184); parity unpinned there, checked only by MSC32<->MSC64 invariants.
 */
#include "pa_oracle.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int EQT = 128;    // EQthreads, SSC:62 / MSC:66
constexpr int NT = 1024;    // Nthreads,  SSC:63 / MSC:67
constexpr int NBMAX = 10000;  // nBmax, SSC:57
constexpr uint32_t LCG_A = 1664525u;     // MSC:53
constexpr uint32_t LCG_C = 1013904223u;  // MSC:54

/* ---------------- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---------------- */
inline void philox10(const uint32_t c_in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
  uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* One per-thread generator as created by curand_init(seed, subsequence, 0)
 * (SSC:152,164,241; MSC:171,195,355).  Word p of subsequence s is output word p%4 of
 * Philox(counter = {p/4 lo, p/4 hi, s lo, s hi}, key = {seed lo, seed hi}), words taken
 * in x,y,z,w order (layout as in rocrand_philox4x32_10.h:124-137,170-186,212-219). */
struct Stream {
  uint32_t k0, k1, s0, s1;
  uint64_t blk;
  uint32_t out[4];
  int sub;
  Stream() {}
  Stream(uint64_t seed, uint64_t subseq) { reset(seed, subseq); }
  void reset(uint64_t seed, uint64_t subseq) {
    k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
    s0 = (uint32_t)subseq; s1 = (uint32_t)(subseq >> 32);
    blk = 0; sub = 0; gen();
  }
  void gen() {
    uint32_t c[4] = {(uint32_t)blk, (uint32_t)(blk >> 32), s0, s1};
    philox10(c, k0, k1, out);
  }
  uint32_t next() {  // curand()
    uint32_t r = out[sub++];
    if (sub == 4) { sub = 0; ++blk; gen(); }
    return r;
  }
};

inline double uniform_map(uint32_t x, uint32_t y, int umap) {
  const double inv53 = 1.0 / 9007199254740992.0;  // 2^-53
  switch (umap) {
    case ORA_U_ROCRAND_UNIFORM: {
      uint64_t z = (uint64_t)x | ((uint64_t)(y >> 11) << 32);
      return inv53 + (double)z * inv53;
    }
    case ORA_U_XOR21_P1: {
      uint64_t z = (uint64_t)x ^ ((uint64_t)y << 21);
      return inv53 + (double)z * inv53;
    }
    case ORA_U_XOR21_HALF: {
      uint64_t z = (uint64_t)x ^ ((uint64_t)y << 21);
      return (double)z * inv53 + inv53 * 0.5;
    }
    case ORA_U_XOR21_SWAP: {
      uint64_t z = (uint64_t)y ^ ((uint64_t)x << 21);
      return inv53 + (double)z * inv53;
    }
    case ORA_U_U64_XHI: {
      uint64_t v = ((uint64_t)x << 32) | y;
      return inv53 + (double)(v >> 11) * inv53;
    }
    case ORA_U_U64_YHI: {
      uint64_t v = ((uint64_t)y << 32) | x;
      return inv53 + (double)(v >> 11) * inv53;
    }
  }
  return 0.0;
}

inline double uniform32_map(uint32_t v, int umap) {
  const double inv32 = 1.0 / 4294967296.0;  // 2^-32
  switch (umap) {
    case ORA_U32_OPEN_CLOSED: return inv32 + (double)v * inv32;
    case ORA_U32_MIDPOINT: return (double)v * inv32 + inv32 * 0.5;
    case ORA_U32_CLOSED_OPEN: return (double)v * inv32;
  }
  return 0.0;
}

/* curand_uniform_double (SSC:154-155,175,187,245; MSC:356).  Single-word mappings draw one
 * curand() word, the two-word candidates draw two. */
inline double stream_uniform(Stream& s, int umap) {
  if (umap >= ORA_U32_OPEN_CLOSED) return uniform32_map(s.next(), umap);
  uint32_t x = s.next();
  uint32_t y = s.next();
  return uniform_map(x, y, umap);
}

/* ---------------- FP64 exp and log ----------------
 * exp: Cody & Waite (1980) argument reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split
 * in two doubles (the high part product exact under fma), Taylor polynomial of degree 13
 * by Horner's rule, exp(x) = 2^k (1 + (r + r^2 P(r))).
 * log: x = 2^e m, m in (sqrt(1/2), sqrt(2)], f = m-1, s = f/(2+f),
 * log(1+f) = f - (f^2/2 - s (f^2/2 + R(s^2))), R(z) = sum_{k=1..11} 2 z^k/(2k+1)
 * (the atanh series, Abramowitz & Stegun 4.1.27), log x = e ln2 + log(1+f) with
 * ln2 split so that e*ln2_hi is exact.
 * The engine implements the same operation sequence, so both agree bit for bit. */
inline double u64_as_double(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }
inline uint64_t double_as_u64(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }

double my_exp(double x) {
  if (x != x) return x;
  if (x > 709.79) return HUGE_VAL;
  if (x < -745.2) return 0.0;
  const double INV_LN2 = 1.4426950408889634;
  const double LN2_HI = 0.6931471805599453;
  const double LN2_LO = 2.3190468138462996e-17;
  double k = std::nearbyint(x * INV_LN2);
  double r = std::fma(-k, LN2_HI, x);
  r = std::fma(-k, LN2_LO, r);
  double p = 1.6059043836821613e-10;
  p = std::fma(p, r, 2.08767569878681e-09);
  p = std::fma(p, r, 2.505210838544172e-08);
  p = std::fma(p, r, 2.755731922398589e-07);
  p = std::fma(p, r, 2.7557319223985893e-06);
  p = std::fma(p, r, 2.48015873015873e-05);
  p = std::fma(p, r, 0.0001984126984126984);
  p = std::fma(p, r, 0.001388888888888889);
  p = std::fma(p, r, 0.008333333333333333);
  p = std::fma(p, r, 0.041666666666666664);
  p = std::fma(p, r, 0.16666666666666666);
  p = std::fma(p, r, 0.5);
  double rr = r * r;
  double t = std::fma(rr, p, r);
  double y = 1.0 + t;
  int ki = (int)k;
  if (ki > 1023) { y *= 2.0; ki -= 1; }
  if (ki < -1022) {
    y *= u64_as_double((uint64_t)(ki + 600 + 1023) << 52);
    y *= u64_as_double((uint64_t)(-600 + 1023) << 52);
    return y;
  }
  return y * u64_as_double((uint64_t)(ki + 1023) << 52);
}

double my_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return std::nan("");
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int e = 0;
  if (x < 2.2250738585072014e-308) { x *= 18014398509481984.0; e = -54; }  // 2^54
  uint64_t b = double_as_u64(x);
  e += (int)(b >> 52) - 1023;
  double m = u64_as_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) { m *= 0.5; e += 1; }
  const double LN2_HI = 0.6931471803691238;   // ln2 with 21 low mantissa bits cleared
  const double LN2_LO = 1.9082149292705877e-10;
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
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
  double hfsq = 0.5 * f * f;
  double dk = (double)e;
  return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
}

/* ---------------- reference reduction tree ----------------
 * warpReduceSum (SSC:107-111): lane i <- v[i] + v[i+off], off = 16,8,4,2,1 (lane 0 kept).
 * blockReduceSum (SSC:113-124): per-warp results, then warp 0 reduces them (zeros for
 * lanes >= blockDim/32). */
inline double warp_tree(double* v) {
  for (int off = 16; off > 0; off >>= 1)
    for (int i = 0; i < off; ++i) v[i] = v[i] + v[i + off];
  return v[0];
}

double block_reduce(const double* vals, int n) {  // n = blockDim.x, multiple of 32
  double w[32];
  int nw = n / 32;
  for (int k = 0; k < 32; ++k) w[k] = 0.0;
  for (int k = 0; k < nw; ++k) {
    double v[32];
    for (int i = 0; i < 32; ++i) v[i] = vals[32 * k + i];
    w[k] = warp_tree(v);
  }
  return warp_tree(w);
}

/* smallblockReduceSum MSC:96-105 (blocks of fewer than 32 threads, power of 2): shared[t]
 * += shared[t + stride] for stride = n/2, n/4, ..., 1; the result is shared[0]. */
double small_block_reduce(const double* vals, int n) {
  double s[32];
  for (int i = 0; i < n; ++i) s[i] = vals[i];
  for (int stride = n >> 1; stride > 0; stride >>= 1)
    for (int t = 0; t < stride; ++t) s[t] += s[t + stride];
  return s[0];
}

inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser (Steele, Lea, Flood 2014)
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Order-independent hash over (index, 64-bit chunk) pairs: sum_i mix64(mix64(i) ^ v_i).
 * The engine computes the same sum with a device reduction. */
uint64_t hash_chunks(const uint8_t* data, uint64_t nbytes_per_item, uint64_t items) {
  uint64_t chunks = (nbytes_per_item + 7) / 8;
  uint64_t h = 0;
#pragma omp parallel for reduction(+ : h) schedule(static)
  for (int64_t it = 0; it < (int64_t)items; ++it) {
    const uint8_t* p = data + (uint64_t)it * nbytes_per_item;
    for (uint64_t c = 0; c < chunks; ++c) {
      uint64_t v = 0;
      uint64_t nb = nbytes_per_item - c * 8 < 8 ? nbytes_per_item - c * 8 : 8;
      memcpy(&v, p + c * 8, nb);
      h += mix64(mix64((uint64_t)it * chunks + c) ^ v);
    }
  }
  return h;
}

uint64_t hash_u32(const uint32_t* v, uint64_t n) {
  uint64_t h = 0;
  for (uint64_t i = 0; i < n; ++i) h += mix64(mix64(i) ^ (uint64_t)v[i]);
  return h;
}

/* ---------------- geometry (SSC:168-172, 180-184; energy SSC:202-207) ----------------
 * Sublattice site idx = ty*L/2 + tx; A site at x = 2tx + (ty&1), B site at 2tx + 1-(ty&1). */
struct Geo {
  int L, Ld2, N, Nh;
  explicit Geo(int L_) : L(L_), Ld2(L_ / 2), N(L_ * L_), Nh(L_ * L_ / 2) {}
  // neighbours in the other sublattice of a site of sublattice A (side=0) or B (side=1)
  inline void nbrs(int side, int idx, int& i1, int& i3, int& i4) const {
    int ty = idx / Ld2, tx = idx - ty * Ld2;
    if (side == 0) i1 = ty * Ld2 + ((ty & 1) ? (tx + 1) : (tx + Ld2 - 1)) % Ld2;
    else i1 = ty * Ld2 + ((ty & 1) ? (tx + Ld2 - 1) : (tx + 1)) % Ld2;
    i3 = ((ty + L - 1) % L) * Ld2 + tx;
    i4 = ((ty + 1) % L) * Ld2 + tx;
  }
  // nb[side][idx] = {i1, i3, i4}
  std::vector<int> table() const {
    std::vector<int> t((size_t)2 * Nh * 3);
    for (int side = 0; side < 2; ++side)
      for (int idx = 0; idx < Nh; ++idx)
        nbrs(side, idx, t[((size_t)side * Nh + idx) * 3], t[((size_t)side * Nh + idx) * 3 + 1],
             t[((size_t)side * Nh + idx) * 3 + 2]);
    return t;
  }
};

/* ---------------- SSC stages ---------------- */

/* ReplicaInit SSC:150-157: per virtual thread t of block B, stream IS + t + 128 B. */
void ssc_init(const Geo& g, int R, uint64_t seed, uint64_t IS, int umap, int8_t* spins) {
#pragma omp parallel for schedule(static)
  for (int B = 0; B < R; ++B) {
    int8_t* gA = spins + (size_t)B * g.N;
    int8_t* gB = gA + g.Nh;
    for (int t = 0; t < EQT; ++t) {
      Stream s(seed, IS + (uint64_t)t + (uint64_t)B * EQT);
      for (int idx = t; idx < g.Nh; idx += EQT) {
        gA[idx] = (stream_uniform(s, umap) >= 0.5) ? 1 : -1;
        gB[idx] = (stream_uniform(s, umap) >= 0.5) ? 1 : -1;
      }
    }
  }
}

/* Boltzmann table SSC:512-516 (host libm exp). */
void ssc_table(double beta, double T[10]) {
  for (int i0 = 0; i0 <= 1; ++i0)
    for (int i1 = 0; i1 <= 4; ++i1)
      T[i0 + 2 * i1] = std::min(1.0, std::exp(-2 * (2 * i0 - 1) * (2 * i1 - 4) * beta));
}

/* checkKerALL SSC:159-193. */
void ssc_sweep(const Geo& g, int R, int8_t* spins, uint32_t sweeps, uint64_t seed, uint64_t IS,
               const double T[10], int umap) {
  const std::vector<int> nbt = g.table();
#pragma omp parallel for schedule(static)
  for (int B = 0; B < R; ++B) {
    int8_t* gA = spins + (size_t)B * g.N;
    int8_t* gB = gA + g.Nh;
    Stream st[EQT];
    for (int t = 0; t < EQT; ++t) st[t].reset(seed, IS + (uint64_t)t + (uint64_t)B * EQT);
    for (uint32_t sw = 0; sw < sweeps; ++sw) {
      for (int side = 0; side < 2; ++side) {
        int8_t* own = side ? gB : gA;
        const int8_t* oth = side ? gA : gB;
        for (int t = 0; t < EQT; ++t) {
          for (int idx = t; idx < g.Nh; idx += EQT) {
            const int* nb = &nbt[((size_t)side * g.Nh + idx) * 3];
            int i1 = nb[0], i3 = nb[1], i4 = nb[2];
            int spin = own[idx];
            int i0 = (1 + spin) / 2;
            int i1t = (4 + oth[i1] + oth[idx] + oth[i3] + oth[i4]) / 2;
            if (stream_uniform(st[t], umap) < T[i0 + 2 * i1t]) own[idx] = (int8_t)-spin;
          }
        }
      }
    }
  }
}

/* energyKer SSC:195-215: left and upper bond of every site, E = -sum s s'. */
void ssc_energy(const Geo& g, int R, const int8_t* spins, int32_t* IE, int32_t* M) {
#pragma omp parallel for schedule(static)
  for (int B = 0; B < R; ++B) {
    const int8_t* gA = spins + (size_t)B * g.N;
    const int8_t* gB = gA + g.Nh;
    int E = 0, Mg = 0;
    for (int idx = 0; idx < g.Nh; ++idx) {
      int ty = idx / g.Ld2, tx = idx - ty * g.Ld2;
      int iL = ty * g.Ld2 + (tx + g.Ld2 - 1) % g.Ld2;
      int iU = ((ty + 1) % g.L) * g.Ld2 + tx;
      int sA = gA[idx], sB = gB[idx], Ai2, Bi2;
      if (ty & 1) { Ai2 = sB; Bi2 = gA[iL]; }
      else { Ai2 = gB[iL]; Bi2 = sA; }
      int Ai4 = gB[iU], Bi4 = gA[iU];
      E += -sA * (Ai2 + Ai4) - sB * (Bi2 + Bi4);
      Mg += sA + sB;
    }
    IE[B] = E;
    M[B] = Mg;
  }
}

/* QKer SSC:217-234: block sums over 1024 replicas, then chunked block sum of those. */
double ssc_q(int R, const int32_t* IE, double dB, double Emean) {
  int nb = (R + NT - 1) / NT;
  std::vector<double> bsum(nb);
  std::vector<double> v(NT);
  for (int b = 0; b < nb; ++b) {
    for (int t = 0; t < NT; ++t) {
      int idx = t + NT * b;
      v[t] = (idx < R) ? my_exp(-dB * ((double)IE[idx] - Emean)) : 0.0;
    }
    bsum[b] = block_reduce(v.data(), NT);
  }
  double acc = 0.0;
  for (int j = 0; (int64_t)j * NT < R; j += NT) {
    for (int t = 0; t < NT; ++t) v[t] = (t + j < nb) ? bsum[t + j] : 0.0;
    acc += block_reduce(v.data(), NT);
  }
  return acc;
}

/* Host lnQ, SSC:533 / MSC:696 (log from the own routine, see my_log). */
double lnq(double dB, double Emean, double q, int Rprev) {
  return -(dB)*Emean + my_log(q) - my_log((double)Rprev);
}

/* CalcTauKer SSC:236-254; exponent a*b+c contracted to fma as nvcc does by default. */
int64_t ssc_tau(int R, const int32_t* IE, int Rinit, double dB, double lnQ, uint64_t seed,
                uint64_t IS, int umap, uint32_t* Roff) {
  double ratio = (double)Rinit / R;
  int64_t tot = 0;
#pragma omp parallel for reduction(+ : tot) schedule(static)
  for (int idx = 0; idx < R; ++idx) {
    Stream s(seed, IS + (uint64_t)idx);
    double mu = ratio * my_exp(std::fma(-dB, (double)IE[idx], -lnQ));
    double mf = std::floor(mu);
    uint32_t r = (stream_uniform(s, umap) < (mu - mf)) ? (uint32_t)(mf + 1) : (uint32_t)mf;
    Roff[idx] = r;
    tot += r;
  }
  return tot;
}

/* CalcParSum SSC:256-270 (exclusive prefix) + resampleKer SSC:272-280 (all N sites copied;
 * the reference grid N/2/128 drops N/2 mod 128 sites when L is not a multiple of 16). */
void ssc_resample(const Geo& g, int Rold, const int8_t* spins, const uint32_t* Roff,
                  int8_t* spins_new) {
  std::vector<uint64_t> pre(Rold);
  uint64_t acc = 0;
  for (int b = 0; b < Rold; ++b) { pre[b] = acc; acc += Roff[b]; }
#pragma omp parallel for schedule(dynamic, 64)
  for (int b = 0; b < Rold; ++b)
    for (uint32_t j = 0; j < Roff[b]; ++j)
      memcpy(spins_new + (pre[b] + j) * (size_t)g.N, spins + (size_t)b * g.N, g.N);
}

/* CalcAverages SSC:282-292: 5 block tree sums, combined in block order. */
void ssc_averages(int R, const int32_t* IE, const int32_t* M, double out[5]) {
  int nb = (R + NT - 1) / NT;
  for (int q = 0; q < 5; ++q) out[q] = 0.0;
  std::vector<double> v[5];
  for (int q = 0; q < 5; ++q) v[q].resize(NT);
  for (int b = 0; b < nb; ++b) {
    for (int t = 0; t < NT; ++t) {
      int idx = t + NT * b;
      double e = 0, m = 0;
      if (idx < R) { e = IE[idx]; m = M[idx]; if (m < 0) m = -m; }
      double e2 = e * e, m2 = m * m, m4 = m2 * m2;
      v[0][t] = e; v[1][t] = e2; v[2][t] = m; v[3][t] = m2; v[4][t] = m4;
    }
    for (int q = 0; q < 5; ++q) out[q] += block_reduce(v[q].data(), NT);
  }
}

/* ---------------- MSC stages (W = uint8_t, uint16_t, uint32_t, uint64_t: MSbits 8/16/32/64) ---------------- */

/* ReplicaInit MSC:169-188.  8 and 16 bits keep the low byte / half of one curand() word
 * (MSC:173-178), 32 bits the whole word.  For 64 bits the two curand() calls of MSC:183-184
 * have no specified order; the first word drawn becomes the high half (decision, unpinned). */
template <typename W>
void msc_init(const Geo& g, int rg, int R, uint64_t seed, uint64_t IS, W* words, uint8_t* active) {
  const int w = (int)sizeof(W) * 8;
#pragma omp parallel for schedule(static)
  for (int B = 0; B < rg; ++B) {
    W* gA = words + (size_t)B * g.N;
    W* gB = gA + g.Nh;
    for (int t = 0; t < EQT; ++t) {
      Stream s(seed, IS + (uint64_t)t + (uint64_t)B * EQT);
      for (int idx = t; idx < g.Nh; idx += EQT) {
        if (w <= 32) {
          gA[idx] = (W)s.next();  // & 0xFF, & 0xFFFF or the whole word
          gB[idx] = (W)s.next();
        } else {
          uint64_t hi = s.next(); uint64_t lo = s.next();
          gA[idx] = (W)((hi << 32) + lo);
          hi = s.next(); lo = s.next();
          gB[idx] = (W)((hi << 32) + lo);
        }
      }
    }
    for (int t = 0; t < w; ++t) active[(size_t)B * w + t] = ((int64_t)B * w + t < R) ? 1 : 0;
  }
}

void msc_table(double beta, uint32_t T[2]) {  // MSC:674-675 (host libm exp)
  T[0] = (uint32_t)std::ceil(4294967296. * std::exp(-4 * beta));
  T[1] = (uint32_t)std::ceil(4294967296. * std::exp(-8 * beta));
}

/* checkKerALL MSC:190-266: bit-sliced count of antiparallel neighbours, one curand() word
 * per site word, expanded to one 32-bit LCG value per replica bit. */
template <typename W>
void msc_sweep(const Geo& g, int rg, W* words, uint32_t sweeps, uint64_t seed, uint64_t IS,
               const uint32_t T[2]) {
  const int w = (int)sizeof(W) * 8;
  const uint32_t T4 = T[0], T8 = T[1];
  const std::vector<int> nbt = g.table();
#pragma omp parallel for schedule(static)
  for (int B = 0; B < rg; ++B) {
    W* gA = words + (size_t)B * g.N;
    W* gB = gA + g.Nh;
    Stream st[EQT];
    for (int t = 0; t < EQT; ++t) st[t].reset(seed, IS + (uint64_t)t + (uint64_t)B * EQT);
    for (uint32_t sw = 0; sw < sweeps; ++sw) {
      for (int side = 0; side < 2; ++side) {
        W* own = side ? gB : gA;
        const W* oth = side ? gA : gB;
        for (int t = 0; t < EQT; ++t) {
          for (int idx = t; idx < g.Nh; idx += EQT) {
            const int* nb = &nbt[((size_t)side * g.Nh + idx) * 3];
            int i1 = nb[0], i3 = nb[1], i4 = nb[2];
            W s = own[idx];
            W I1 = s ^ oth[i1], I2 = s ^ oth[idx], I3 = s ^ oth[i3], I4 = s ^ oth[i4];
            W x12 = I1 ^ I2, x34 = I3 ^ I4, a12 = I1 & I2, a34 = I3 & I4;
            W sum0 = x12 ^ x34;
            W sum1 = (x12 & x34) ^ a12 ^ a34;
            W sum2 = a12 & a34;
            W c4 = 0, c8 = 0;
            uint32_t ran = st[t].next();
            for (int i = 0; i < w; ++i) {
              c4 |= (W)(ran < T4) << i;
              c8 |= (W)(ran < T8) << i;
              ran = LCG_A * ran + LCG_C;
            }
            W s12 = sum1 | sum2;
            W acc = s12 | (~s12 & ((sum0 & c4) | (~sum0 & c8)));
            own[idx] = s ^ acc;
          }
        }
      }
    }
  }
}

/* energyKer MSC:268-317: e = 2*count - 4 per site pair, m = 2(bA + bB) - 2. */
template <typename W>
void msc_energy(const Geo& g, int rg, const W* words, int32_t* IE, int32_t* M) {
  const int w = (int)sizeof(W) * 8;
#pragma omp parallel for schedule(static)
  for (int B = 0; B < rg; ++B) {
    const W* gA = words + (size_t)B * g.N;
    const W* gB = gA + g.Nh;
    std::vector<int> E(w, 0), Mg(w, 0);
    for (int idx = 0; idx < g.Nh; ++idx) {
      int ty = idx / g.Ld2, tx = idx - ty * g.Ld2;
      int iL = ty * g.Ld2 + (tx + g.Ld2 - 1) % g.Ld2;
      int iU = ((ty + 1) % g.L) * g.Ld2 + tx;
      W sA = gA[idx], sB = gB[idx], Ai2, Bi2;
      if (ty & 1) { Ai2 = sB; Bi2 = gA[iL]; }
      else { Ai2 = gB[iL]; Bi2 = sA; }
      W Ai4 = gB[iU], Bi4 = gA[iU];
      W I1 = sA ^ Ai2, I2 = sA ^ Ai4, I3 = sB ^ Bi2, I4 = sB ^ Bi4;
      W x12 = I1 ^ I2, x34 = I3 ^ I4, a12 = I1 & I2, a34 = I3 & I4;
      W sum0 = x12 ^ x34, sum1 = (x12 & x34) ^ a12 ^ a34, sum2 = a12 & a34;
      for (int i = 0; i < w; ++i) {
        E[i] += 2 * ((int)((sum0 >> i) & 1) + 2 * (int)((sum1 >> i) & 1) + 4 * (int)((sum2 >> i) & 1)) - 4;
        Mg[i] += 2 * ((int)((sA >> i) & 1) + (int)((sB >> i) & 1)) - 2;
      }
    }
    for (int i = 0; i < w; ++i) { IE[(size_t)B * w + i] = E[i]; M[(size_t)B * w + i] = Mg[i]; }
  }
}

/* QKer MSC:319-345: lane sums per group (block of w threads: blockReduceSum for 32 and 64,
 * smallblockReduceSum for 8 and 16, MSC:324-328), 1024-group block sums, then the chunked
 * final sum. */
double msc_q(int w, int rg, const uint8_t* active, const int32_t* IE, double dB, double Emean) {
  std::vector<double> gsum(rg);
  for (int gi = 0; gi < rg; ++gi) {
    double v[64];
    for (int br = 0; br < w; ++br) {
      size_t k = (size_t)gi * w + br;
      v[br] = active[k] ? my_exp(-dB * ((double)IE[k] - Emean)) : 0.0;
    }
    gsum[gi] = w < 32 ? small_block_reduce(v, w) : block_reduce(v, w);
  }
  int nb = (rg + NT - 1) / NT;
  std::vector<double> bsum(nb), v(NT);
  for (int b = 0; b < nb; ++b) {
    for (int t = 0; t < NT; ++t) { int idx = t + NT * b; v[t] = (idx < rg) ? gsum[idx] : 0.0; }
    bsum[b] = block_reduce(v.data(), NT);
  }
  double acc = 0.0;
  for (int j = 0; (int64_t)j * NT < rg; j += NT) {
    for (int t = 0; t < NT; ++t) v[t] = (t + j < nb) ? bsum[t + j] : 0.0;
    acc += block_reduce(v.data(), NT);
  }
  return acc;
}

/* CalcTauKer MSC:347-360: stream IS + br + w*idx for each active lane. */
int64_t msc_tau(int w, int rg, const uint8_t* active, const int32_t* IE, int Rinit, int R,
                double dB, double lnQ, uint64_t seed, uint64_t IS, int umap, uint32_t* Roff) {
  double ratio = (double)Rinit / R;
  int64_t tot = 0;
#pragma omp parallel for reduction(+ : tot) schedule(static)
  for (int idx = 0; idx < rg; ++idx) {
    for (int br = 0; br < w; ++br) {
      size_t k = (size_t)idx * w + br;
      if (!active[k]) { Roff[k] = 0; continue; }
      double mu = ratio * my_exp(std::fma(-dB, (double)IE[k], -lnQ));
      double mf = std::floor(mu);
      Stream s(seed, IS + (uint64_t)br + (uint64_t)w * idx);
      uint32_t r = (stream_uniform(s, umap) < (mu - mf)) ? (uint32_t)(mf + 1) : (uint32_t)mf;
      Roff[k] = r;
      tot += r;
    }
  }
  return tot;
}

/* CalcParSum MSC:362-395 (prefix in group-major, lane-minor order) and resampleKer
 * MSC:397-435: child n = prefix + p goes to group n mod rg_new, lane n div rg_new; the new
 * population starts zeroed (MSC:720); isActive and IE are set for every child. */
template <typename W>
void msc_resample(const Geo& g, int rg_old, const W* words, const uint8_t* active,
                  const int32_t* IE, const uint32_t* Roff, int Rnew, W* words_new,
                  uint8_t* active_new, int32_t* IE_new) {
  const int w = (int)sizeof(W) * 8;
  int rg_new = (int)std::ceil(Rnew / (float)w);
  memset(words_new, 0, (size_t)rg_new * g.N * sizeof(W));
  memset(active_new, 0, (size_t)rg_new * w);
  memset(IE_new, 0, (size_t)rg_new * w * sizeof(int32_t));
  // parent of every child in population order
  std::vector<int64_t> parent(Rnew);
  {
    int64_t n = 0;
    for (int64_t k = 0; k < (int64_t)rg_old * w; ++k)
      for (uint32_t p = 0; p < Roff[k]; ++p) parent[n++] = k;
  }
  for (int64_t n = 0; n < Rnew; ++n) {
    int kk = (int)(n / rg_new), jj = (int)(n % rg_new);
    active_new[(size_t)jj * w + kk] = 1;
    IE_new[(size_t)jj * w + kk] = IE[parent[n]];
  }
#pragma omp parallel for schedule(static)
  for (int jj = 0; jj < rg_new; ++jj) {
    W* dst = words_new + (size_t)jj * g.N;
    for (int kk = 0; kk < w; ++kk) {
      int64_t n = (int64_t)kk * rg_new + jj;
      if (n >= Rnew) continue;
      int64_t par = parent[n];
      int pg = (int)(par / w), pb = (int)(par % w);
      const W* src = words + (size_t)pg * g.N;
      for (int s = 0; s < g.N; ++s) dst[s] |= (W)((src[s] >> pb) & 1) << kk;
    }
  }
}

/* CalcAverages MSC:437-450: grid (blocks of 1024 groups) x (w lanes); block sums are
 * combined in launch order (lane-major over blocks). */
void msc_averages(int w, int rg, const uint8_t* active, const int32_t* IE, const int32_t* M,
                  double out[5]) {
  int nb = (rg + NT - 1) / NT;
  for (int q = 0; q < 5; ++q) out[q] = 0.0;
  std::vector<double> v[5];
  for (int q = 0; q < 5; ++q) v[q].resize(NT);
  for (int by = 0; by < w; ++by) {
    for (int b = 0; b < nb; ++b) {
      for (int t = 0; t < NT; ++t) {
        int idx = t + NT * b;
        double e = 0, m = 0;
        if (idx < rg && active[(size_t)idx * w + by]) {
          e = IE[(size_t)idx * w + by]; m = M[(size_t)idx * w + by]; if (m < 0) m = -m;
        }
        double e2 = e * e, m2 = m * m, m4 = m2 * m2;
        v[0][t] = e; v[1][t] = e2; v[2][t] = m; v[3][t] = m2; v[4][t] = m4;
      }
      for (int q = 0; q < 5; ++q) out[q] += block_reduce(v[q].data(), NT);
    }
  }
}

void set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* HistogramOverlap + CalcOverlap, SSC:308-334: q and lnQ of the candidate step as in the
 * step itself (QKer, SSC:533), then sum over replicas of min(1, Rinit/R exp(-dB E - lnQ))
 * with the block tree; the reference adds the block sums with atomicAdd (order unspecified),
 * here they are added in launch order.  Returns the sum divided by R. */
double ssc_overlap(int R, const int32_t* IE, int Rinit, double dB, double Emean) {
  const double q = ssc_q(R, IE, dB, Emean);
  const double lnQ = lnq(dB, Emean, q, R);
  const double ratio = (double)Rinit / R;
  int nb = (R + NT - 1) / NT;
  std::vector<double> v(NT);
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) {
    for (int t = 0; t < NT; ++t) {
      int idx = t + NT * b;
      v[t] = (idx < R) ? std::min(1.0, ratio * my_exp(std::fma(-dB, (double)IE[idx], -lnQ))) : 0.0;
    }
    acc += block_reduce(v.data(), NT);
  }
  return acc / R;
}

/* MSC:466-499: the same over active lanes; grid (blocks of 1024 groups) x (w lanes), block
 * sums added in launch order (lane-major over blocks), as for the averages. */
double msc_overlap(int w, int rg, const uint8_t* active, const int32_t* IE, int Rinit, int R,
                   double dB, double Emean) {
  const double q = msc_q(w, rg, active, IE, dB, Emean);
  const double lnQ = lnq(dB, Emean, q, R);
  const double ratio = (double)Rinit / R;
  int nb = (rg + NT - 1) / NT;
  std::vector<double> v(NT);
  double acc = 0.0;
  for (int by = 0; by < w; ++by) {
    for (int b = 0; b < nb; ++b) {
      for (int t = 0; t < NT; ++t) {
        int idx = t + NT * b;
        size_t k = (size_t)idx * w + by;
        v[t] = (idx < rg && active[k])
                   ? std::min(1.0, ratio * my_exp(std::fma(-dB, (double)IE[k], -lnQ)))
                   : 0.0;
      }
      acc += block_reduce(v.data(), NT);
    }
  }
  return acc / R;
}

/* Next inverse-temperature step of the adaptive schedule, SSC:644-660 / MSC:795-811: grow
 * the step by 10 % while the overlap stays >= MaxOverlap (and B + step < Bfin), then bisect
 * on [0, step] until MinOverlap <= overlap < MaxOverlap; the last step lands on Bfin.
 * `ovl(dB)` evaluates CalcOverlap on the current population. */
template <typename F>
int adaptive_step(const ora_params* p, double Bprev, double& deltaBeta, F ovl) {
  double overlap, dBmin = 0, dBmax = deltaBeta, dBmean;
  for (int it = 0;; ++it) {
    if (it > 100000) return -1;
    overlap = ovl(dBmax);
    if ((overlap >= p->max_overlap) && (Bprev + dBmax < p->Bfin)) dBmax *= 1.1;
    else break;
  }
  if (overlap >= p->min_overlap) {
    dBmean = dBmax;
  } else {
    for (int it = 0;; ++it) {
      if (it > 100000) return -1;
      dBmean = 0.5 * (dBmin + dBmax);
      overlap = ovl(dBmean);
      if (overlap < p->min_overlap) dBmax = dBmean;
      else if (overlap >= p->max_overlap) dBmin = dBmean;
      else break;
    }
  }
  if ((Bprev < p->Bfin) && (Bprev + dBmean > p->Bfin)) deltaBeta = p->Bfin - Bprev;
  else deltaBeta = dBmean;
  return 0;
}

/* UpdateShistE, SSC:296-302 / MSC:454-460: histogram bin (2N + E)/4 per replica. */
void hist_add(int N, int64_t n, const int32_t* IE, const uint8_t* active, std::vector<int32_t>& H) {
  for (int64_t k = 0; k < n; ++k)
    if (!active || active[k]) H[(2 * N + IE[k]) / 4] += 1;
}

/* Multi-histogram reweighting, SSC:705-764 (MSC:855-914).  Ei[k] = 4k - 2N (SSC:461-463).
 * out: nB rows of 1 + 3 niter doubles (B, then E/N, C/N, BF/N per iteration).  BF is
 * overwritten with the reweighted free energies of each iteration, as in the reference. */
void mhr_analysis(int N, int nB, const double* B, double* BF, const int32_t* R,
                  const int32_t* hist, int niter, double* out) {
  const int K = N + 1, cols = 1 + 3 * niter;
  std::vector<int> Ei(K);
  for (int k = 0; k < K; ++k) Ei[k] = 4 * k - 2 * N;
  std::vector<double> lnOmega(K), Sigma(nB), expOm(K), Om(K);
  std::vector<bool> relTerm(K);
  for (int i = 0; i < nB; ++i) out[(size_t)i * cols] = B[i];
  for (int l = 0; l < niter; ++l) {
    for (int k = 0; k < K; ++k) {
      Sigma[0] = BF[0] - B[0] * Ei[k];
      double mSigma = Sigma[0];
      for (int i = 1; i < nB; ++i) {
        Sigma[i] = BF[i] - B[i] * Ei[k];
        if (mSigma < Sigma[i]) mSigma = Sigma[i];
      }
      double sD = 0;
      for (int i = 0; i < nB; ++i) sD += R[i] * exp(Sigma[i] - mSigma);
      if ((hist[k] == 0) || (sD == 0)) {
        relTerm[k] = false;
        lnOmega[k] = 0;
      } else {
        relTerm[k] = true;
        lnOmega[k] = log(hist[k]) - mSigma - log(sD);
      }
    }
    for (int i = 0; i < nB; ++i) {
      double mOm = lnOmega[0] - B[i] * Ei[0];
      for (int k = 0; k < K; ++k) {
        Om[k] = lnOmega[k] - B[i] * Ei[k];
        if (mOm < Om[k]) mOm = Om[k];
      }
      double pp = 0;
      for (int k = 0; k < K; ++k) {
        expOm[k] = exp(Om[k] - mOm);
        if (relTerm[k]) pp += expOm[k];
      }
      double s = 0;
      for (int k = 0; k < K; ++k)
        if (relTerm[k]) s += Ei[k] * expOm[k];
      double E_MHR = s / pp / N;
      double BF_MHR = -mOm - log(pp);
      BF[i] = BF_MHR;
      s = 0;
      for (int k = 0; k < K; ++k)
        if (relTerm[k]) s += pow(Ei[k] - E_MHR * N, 2) * expOm[k];
      double C_MHR = B[i] * B[i] * s / pp / N;
      double* o = out + (size_t)i * cols + 1 + 3 * l;
      o[0] = E_MHR;
      o[1] = C_MHR;
      o[2] = BF_MHR / N;
    }
  }
}

struct MhrOut {
  int32_t* hist = nullptr;
  double* mhr = nullptr;
};

/* ---------------- driver: main() of SSC:381-806 / MSC:546-957 ---------------- */
struct Thermo {
  std::vector<double> B, Binc, E, C, M, M2, M4, lnQ, BF, S;
  std::vector<int> R;
  Thermo() : B(NBMAX), Binc(NBMAX), E(NBMAX, 0), C(NBMAX, 0), M(NBMAX, 0), M2(NBMAX, 0),
             M4(NBMAX, 0), lnQ(NBMAX, 0), BF(NBMAX, 0), S(NBMAX, 0), R(NBMAX, 0) {}
};

void fill_row(ora_row& row, const Thermo& th, int i, int N) {
  row.B = th.B[i];
  row.E = th.E[i] / N;
  row.C = th.C[i] / N;
  row.M = th.M[i] / N;
  row.M2 = th.M2[i] / N / N;
  row.M4 = th.M4[i] / N / N / N / N;
  row.BF = th.BF[i] / N;
  row.S = th.S[i] / N;
  row.R = th.R[i];
  row.lnQ = th.lnQ[i];
}

template <typename W>
int run_msc(const ora_params* p, ora_row* rows, int max_rows, int32_t* nrows, double* tot,
            const MhrOut& mo) {
  const int w = (int)sizeof(W) * 8;
  Geo g(p->L);
  const int N = g.N;
  Thermo th;
  th.B[0] = th.Binc[0] = p->Binit;
  th.BF[0] = -N * std::log(2.0);
  th.R[0] = p->Rinit;
  uint64_t IS = p->initial_sequence;
  double totPop = 0;
  for (int r = 0; r < p->runs; ++r) {
    int rg = (int)std::ceil(th.R[0] / (float)w);
    double sumlnQ = 0.0, Emean = 0.0;
    std::vector<W> words((size_t)rg * N), words_new;
    std::vector<uint8_t> active((size_t)rg * w), active_new;
    std::vector<int32_t> IE((size_t)rg * w), M((size_t)rg * w), IE_new;
    std::vector<uint32_t> Roff;
    msc_init<W>(g, rg, th.R[0], p->seed, IS, words.data(), active.data());
    IS += (uint64_t)rg * EQT;
    msc_energy<W>(g, rg, words.data(), IE.data(), M.data());
    std::vector<int32_t> H(p->mhr ? N + 1 : 0, 0);
    if (p->mhr) hist_add(N, (int64_t)rg * w, IE.data(), active.data(), H);  // MSC:660-662
    ora_row* out = rows + (size_t)r * max_rows;
    int i = 1, iprev = 0;
    double deltaBeta = p->dB;
    th.B[i] = th.Binc[i] = th.B[iprev] + deltaBeta;
    int steps = 0;
    while (th.B[i] <= p->Bfin) {
      if (p->max_steps && steps >= p->max_steps) break;
      uint32_t T[2];
      msc_table(th.B[i], T);
      double dB = th.B[i] - th.B[i - 1];
      double q = msc_q(w, rg, active.data(), IE.data(), dB, Emean);
      th.lnQ[i] = lnq(dB, Emean, q, th.R[i - 1]);
      Roff.assign((size_t)rg * w, 0);
      int64_t Rn = msc_tau(w, rg, active.data(), IE.data(), p->Rinit, th.R[i - 1], dB,
                           th.lnQ[i], p->seed, IS, p->umap, Roff.data());
      IS += (uint64_t)rg * w;
      th.R[i] = (int)Rn;
      uint64_t rhash = hash_u32(Roff.data(), Roff.size());
      int rg_new = (int)std::ceil(th.R[i] / (float)w);
      words_new.assign((size_t)rg_new * N, 0);
      active_new.assign((size_t)rg_new * w, 0);
      IE_new.assign((size_t)rg_new * w, 0);
      msc_resample<W>(g, rg, words.data(), active.data(), IE.data(), Roff.data(), th.R[i],
                      words_new.data(), active_new.data(), IE_new.data());
      words.swap(words_new);
      active.swap(active_new);
      IE.swap(IE_new);
      rg = rg_new;
      M.assign((size_t)rg * w, 0);
      msc_sweep<W>(g, rg, words.data(), p->eqsweeps, p->seed, IS, T);
      IS += (uint64_t)rg * EQT;
      msc_energy<W>(g, rg, words.data(), IE.data(), M.data());
      if (p->mhr) hist_add(N, (int64_t)rg * w, IE.data(), active.data(), H);  // MSC:759-762
      double Av[5];
      msc_averages(w, rg, active.data(), IE.data(), M.data(), Av);
      th.E[i] = Emean = Av[0] / th.R[i];
      th.C[i] = (Av[1] / th.R[i] - th.E[i] * th.E[i]) * th.B[i] * th.B[i];
      th.M[i] = Av[2] / th.R[i];
      th.M2[i] = Av[3] / th.R[i];
      th.M4[i] = Av[4] / th.R[i];
      sumlnQ -= th.lnQ[i];
      th.BF[i] = -N * std::log(2.0) + sumlnQ;
      th.S[i] = th.B[i] * th.E[i] - th.BF[i];
      if (i < max_rows) {
        ora_row& row = out[i];
        row.Rprev = th.R[i - 1];
        row.q = q;
        for (int k = 0; k < 5; ++k) row.sums[k] = Av[k];
        int64_t se = 0, sm = 0;
        for (size_t k = 0; k < active.size(); ++k)
          if (active[k]) { se += IE[k]; sm += M[k]; }
        row.sumE_int = se;
        row.sumM_int = sm;
        row.roff_hash = rhash;
        row.spin_hash = hash_chunks((const uint8_t*)words.data(), (uint64_t)N * sizeof(W), rg);
        row.IS_after = IS;
      }
      iprev = i;
      totPop += th.R[i];
      i++;
      steps++;
      if (i >= NBMAX) {
        fprintf(stderr, "Error: number of temperature steps exceeds nBmax=%d.\n", NBMAX);
        return 1;
      }
      if (r == 0) {
        if (p->adaptive &&
            adaptive_step(p, th.B[iprev], deltaBeta, [&](double dBc) {
              return msc_overlap(w, rg, active.data(), IE.data(), p->Rinit, th.R[iprev], dBc,
                                 Emean);
            }))
          return -1;
        th.B[i] = th.Binc[i] = th.B[iprev] + deltaBeta;
      } else {
        th.B[i] = th.Binc[i];
      }
    }
    int nB = i;
    nrows[r] = nB < max_rows ? nB : max_rows;
    for (int k = 0; k < nrows[r]; ++k) {
      ora_row save = out[k];
      fill_row(out[k], th, k, N);
      if (k > 0) {
        out[k].Rprev = save.Rprev; out[k].q = save.q;
        for (int z = 0; z < 5; ++z) out[k].sums[z] = save.sums[z];
        out[k].sumE_int = save.sumE_int; out[k].sumM_int = save.sumM_int;
        out[k].roff_hash = save.roff_hash; out[k].spin_hash = save.spin_hash;
        out[k].IS_after = save.IS_after;
      }
    }
    // row 0 prints the never-assigned stack values as 0 (SSC:430-436, fixtures line 1)
    out[0].E = out[0].C = out[0].M = out[0].M2 = out[0].M4 = out[0].S = out[0].lnQ = 0.0;
    if (p->mhr) {
      // after the results file (SSC:695-791); th.BF[0] keeps the reweighted value for the
      // next run, which prints and uses it
      const int niter = p->mhr_iter > 0 ? p->mhr_iter : 1;
      const int cols = 1 + 3 * niter;
      std::vector<double> mo_rows((size_t)nB * cols);
      mhr_analysis(N, nB, th.B.data(), th.BF.data(), th.R.data(), H.data(), niter,
                   mo_rows.data());
      if (mo.hist) memcpy(mo.hist + (size_t)r * (N + 1), H.data(), (size_t)(N + 1) * 4);
      if (mo.mhr)
        for (int k = 0; k < nB && k < max_rows; ++k)
          memcpy(mo.mhr + ((size_t)r * max_rows + k) * cols, &mo_rows[(size_t)k * cols], cols * 8);
    }
  }
  *tot = totPop;
  return 0;
}

int run_ssc(const ora_params* p, ora_row* rows, int max_rows, int32_t* nrows, double* tot,
            const MhrOut& mo) {
  Geo g(p->L);
  const int N = g.N;
  Thermo th;
  th.B[0] = th.Binc[0] = p->Binit;
  th.BF[0] = -N * std::log(2.0);
  th.R[0] = p->Rinit;
  uint64_t IS = p->initial_sequence;
  double totPop = 0;
  for (int r = 0; r < p->runs; ++r) {
    double sumlnQ = 0.0, Emean = 0.0;
    std::vector<int8_t> spins((size_t)th.R[0] * N), spins_new;
    std::vector<int32_t> IE(th.R[0]), M(th.R[0]);
    std::vector<uint32_t> Roff;
    ssc_init(g, th.R[0], p->seed, IS, p->umap, spins.data());
    IS += (uint64_t)th.R[0] * EQT;
    ssc_energy(g, th.R[0], spins.data(), IE.data(), M.data());
    std::vector<int32_t> H(p->mhr ? N + 1 : 0, 0);
    if (p->mhr) hist_add(N, th.R[0], IE.data(), nullptr, H);  // SSC:496-501
    ora_row* out = rows + (size_t)r * max_rows;
    int i = 1, iprev = 0;
    double deltaBeta = p->dB;
    th.B[i] = th.Binc[i] = th.B[iprev] + deltaBeta;
    int steps = 0;
    while (th.B[i] <= p->Bfin) {
      if (p->max_steps && steps >= p->max_steps) break;
      double T[10];
      ssc_table(th.B[i], T);
      double dB = th.B[i] - th.B[i - 1];
      int Rp = th.R[i - 1];
      double q = ssc_q(Rp, IE.data(), dB, Emean);
      th.lnQ[i] = lnq(dB, Emean, q, Rp);
      Roff.assign(Rp, 0);
      int64_t Rn = ssc_tau(Rp, IE.data(), p->Rinit, dB, th.lnQ[i], p->seed, IS, p->umap,
                           Roff.data());
      IS += (uint64_t)Rp;
      th.R[i] = (int)Rn;
      uint64_t rhash = hash_u32(Roff.data(), Roff.size());
      spins_new.assign((size_t)th.R[i] * N, 0);
      ssc_resample(g, Rp, spins.data(), Roff.data(), spins_new.data());
      spins.swap(spins_new);
      IE.assign(th.R[i], 0);
      M.assign(th.R[i], 0);
      ssc_sweep(g, th.R[i], spins.data(), p->eqsweeps, p->seed, IS, T, p->umap);
      IS += (uint64_t)th.R[i] * EQT;
      ssc_energy(g, th.R[i], spins.data(), IE.data(), M.data());
      if (p->mhr) hist_add(N, th.R[i], IE.data(), nullptr, H);  // SSC:608-611
      double Av[5];
      ssc_averages(th.R[i], IE.data(), M.data(), Av);
      th.E[i] = Emean = Av[0] / th.R[i];
      th.C[i] = (Av[1] / th.R[i] - th.E[i] * th.E[i]) * th.B[i] * th.B[i];
      th.M[i] = Av[2] / th.R[i];
      th.M2[i] = Av[3] / th.R[i];
      th.M4[i] = Av[4] / th.R[i];
      sumlnQ -= th.lnQ[i];
      th.BF[i] = -N * std::log(2.0) + sumlnQ;
      th.S[i] = th.B[i] * th.E[i] - th.BF[i];
      if (i < max_rows) {
        ora_row& row = out[i];
        row.Rprev = Rp;
        row.q = q;
        for (int k = 0; k < 5; ++k) row.sums[k] = Av[k];
        int64_t se = 0, sm = 0;
        for (int k = 0; k < th.R[i]; ++k) { se += IE[k]; sm += M[k]; }
        row.sumE_int = se;
        row.sumM_int = sm;
        row.roff_hash = rhash;
        row.spin_hash = hash_chunks((const uint8_t*)spins.data(), (uint64_t)N, th.R[i]);
        row.IS_after = IS;
      }
      iprev = i;
      totPop += th.R[i];
      i++;
      steps++;
      if (i >= NBMAX) {
        fprintf(stderr, "Error: number of temperature steps exceeds nBmax=%d.\n", NBMAX);
        return 1;
      }
      if (r == 0) {
        if (p->adaptive &&
            adaptive_step(p, th.B[iprev], deltaBeta, [&](double dBc) {
              return ssc_overlap(th.R[iprev], IE.data(), p->Rinit, dBc, Emean);
            }))
          return -1;
        th.B[i] = th.Binc[i] = th.B[iprev] + deltaBeta;
      } else {
        th.B[i] = th.Binc[i];
      }
    }
    int nB = i;
    nrows[r] = nB < max_rows ? nB : max_rows;
    for (int k = 0; k < nrows[r]; ++k) {
      ora_row save = out[k];
      fill_row(out[k], th, k, N);
      if (k > 0) {
        out[k].Rprev = save.Rprev; out[k].q = save.q;
        for (int z = 0; z < 5; ++z) out[k].sums[z] = save.sums[z];
        out[k].sumE_int = save.sumE_int; out[k].sumM_int = save.sumM_int;
        out[k].roff_hash = save.roff_hash; out[k].spin_hash = save.spin_hash;
        out[k].IS_after = save.IS_after;
      }
    }
    out[0].E = out[0].C = out[0].M = out[0].M2 = out[0].M4 = out[0].S = out[0].lnQ = 0.0;
    if (p->mhr) {
      // after the results file (SSC:695-791); th.BF[0] keeps the reweighted value for the
      // next run, which prints and uses it
      const int niter = p->mhr_iter > 0 ? p->mhr_iter : 1;
      const int cols = 1 + 3 * niter;
      std::vector<double> mo_rows((size_t)nB * cols);
      mhr_analysis(N, nB, th.B.data(), th.BF.data(), th.R.data(), H.data(), niter,
                   mo_rows.data());
      if (mo.hist) memcpy(mo.hist + (size_t)r * (N + 1), H.data(), (size_t)(N + 1) * 4);
      if (mo.mhr)
        for (int k = 0; k < nB && k < max_rows; ++k)
          memcpy(mo.mhr + ((size_t)r * max_rows + k) * cols, &mo_rows[(size_t)k * cols], cols * 8);
    }
  }
  *tot = totPop;
  return 0;
}

}  // namespace

/* ---------------- C API ---------------- */
// word-type dispatch of the stage entry points: F is called with a W* null pointer whose type
// selects the instantiation
template <typename F>
void by_width(int w, F f) {
  if (w == 8) f((uint8_t*)nullptr);
  else if (w == 16) f((uint16_t*)nullptr);
  else if (w == 32) f((uint32_t*)nullptr);
  else f((uint64_t*)nullptr);
}

extern "C" {

int ora_run(const ora_params* p, ora_row* rows, int max_rows, int32_t* nrows, double* elapsed_s,
            double* total_pop) {
  return ora_run_ex(p, rows, max_rows, nrows, elapsed_s, total_pop, nullptr, nullptr);
}

void ora_mhr_analysis(int N, int nB, const double* B, double* BF, const int32_t* R,
                      const int32_t* hist, int niter, double* out) {
  mhr_analysis(N, nB, B, BF, R, hist, niter, out);
}

int ora_run_ex(const ora_params* p, ora_row* rows, int max_rows, int32_t* nrows,
               double* elapsed_s, double* total_pop, int32_t* hist, double* mhr) {
  MhrOut mo;
  mo.hist = hist;
  mo.mhr = mhr;
  if (!p || p->L < 2 || (p->L & 1) || p->Rinit < 1 || p->runs < 1) return -1;
  if (p->coding == ORA_MSC && p->msbits != 8 && p->msbits != 16 && p->msbits != 32 &&
      p->msbits != 64)
    return -1;
  ora_params pp = *p;
  if (pp.adaptive && pp.min_overlap == 0.0 && pp.max_overlap == 0.0) {
    pp.min_overlap = 0.85;  // SSC:41
    pp.max_overlap = 0.87;  // SSC:42
  }
  p = &pp;
  set_threads(p->threads);
  memset(rows, 0, sizeof(ora_row) * (size_t)max_rows * p->runs);
  auto t0 = std::chrono::steady_clock::now();
  double tot = 0;
  int rc;
  if (p->coding == ORA_SSC) rc = run_ssc(p, rows, max_rows, nrows, &tot, mo);
  else if (p->msbits == 8) rc = run_msc<uint8_t>(p, rows, max_rows, nrows, &tot, mo);
  else if (p->msbits == 16) rc = run_msc<uint16_t>(p, rows, max_rows, nrows, &tot, mo);
  else if (p->msbits == 32) rc = run_msc<uint32_t>(p, rows, max_rows, nrows, &tot, mo);
  else rc = run_msc<uint64_t>(p, rows, max_rows, nrows, &tot, mo);
  auto t1 = std::chrono::steady_clock::now();
  if (elapsed_s) *elapsed_s = std::chrono::duration<double>(t1 - t0).count();
  if (total_pop) *total_pop = tot;
  return rc;
}

void ora_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox10(ctr, key[0], key[1], out);
}

uint32_t ora_stream_word(uint64_t seed, uint64_t subsequence, uint64_t pos) {
  uint32_t c[4] = {(uint32_t)(pos / 4), (uint32_t)((pos / 4) >> 32), (uint32_t)subsequence,
                   (uint32_t)(subsequence >> 32)};
  uint32_t o[4];
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32), o);
  return o[pos % 4];
}

double ora_uniform(uint32_t x, uint32_t y, int umap) {
  return umap >= ORA_U32_OPEN_CLOSED ? uniform32_map(x, umap) : uniform_map(x, y, umap);
}
double ora_exp(double x) { return my_exp(x); }
double ora_log(double x) { return my_log(x); }
double ora_block_reduce(const double* v, int n) { return block_reduce(v, n); }
double ora_small_block_reduce(const double* v, int n) { return small_block_reduce(v, n); }

uint64_t ora_hash_bytes(uint64_t h, const void* data, uint64_t n) {
  return h + hash_chunks((const uint8_t*)data, n, 1);
}

void ora_ssc_init(int L, int R, uint64_t seed, uint64_t IS, int umap, int8_t* spins) {
  ssc_init(Geo(L), R, seed, IS, umap, spins);
}
void ora_ssc_sweep(int L, int R, int8_t* spins, uint32_t sweeps, uint64_t seed, uint64_t IS,
                   double beta, int umap) {
  double T[10];
  ssc_table(beta, T);
  ssc_sweep(Geo(L), R, spins, sweeps, seed, IS, T, umap);
}
void ora_ssc_energy(int L, int R, const int8_t* spins, int32_t* IE, int32_t* M) {
  ssc_energy(Geo(L), R, spins, IE, M);
}
void ora_msc_init(int L, int w, int rg, int R, uint64_t seed, uint64_t IS, void* words,
                  uint8_t* active) {
  by_width(w, [&](auto* t) {
    using W = std::remove_pointer_t<decltype(t)>;
    msc_init<W>(Geo(L), rg, R, seed, IS, (W*)words, active);
  });
}
void ora_msc_sweep(int L, int w, int rg, void* words, uint32_t sweeps, uint64_t seed,
                   uint64_t IS, double beta) {
  uint32_t T[2];
  msc_table(beta, T);
  by_width(w, [&](auto* t) {
    using W = std::remove_pointer_t<decltype(t)>;
    msc_sweep<W>(Geo(L), rg, (W*)words, sweeps, seed, IS, T);
  });
}
void ora_msc_energy(int L, int w, int rg, const void* words, int32_t* IE, int32_t* M) {
  by_width(w, [&](auto* t) {
    using W = std::remove_pointer_t<decltype(t)>;
    msc_energy<W>(Geo(L), rg, (const W*)words, IE, M);
  });
}
int64_t ora_ssc_tau(int R, const int32_t* IE, int Rinit, double dB, double lnQ, uint64_t seed,
                    uint64_t IS, int umap, uint32_t* Roff) {
  return ssc_tau(R, IE, Rinit, dB, lnQ, seed, IS, umap, Roff);
}
int64_t ora_msc_tau(int w, int rg, const uint8_t* active, const int32_t* IE, int Rinit, int R,
                    double dB, double lnQ, uint64_t seed, uint64_t IS, int umap, uint32_t* Roff) {
  return msc_tau(w, rg, active, IE, Rinit, R, dB, lnQ, seed, IS, umap, Roff);
}
double ora_ssc_q(int R, const int32_t* IE, double dB, double Emean) {
  return ssc_q(R, IE, dB, Emean);
}
double ora_msc_q(int w, int rg, const uint8_t* active, const int32_t* IE, double dB,
                 double Emean) {
  return msc_q(w, rg, active, IE, dB, Emean);
}
double ora_lnq(double dB, double Emean, double q, int Rprev) { return lnq(dB, Emean, q, Rprev); }
double ora_ssc_overlap(int R, const int32_t* IE, int Rinit, double dB, double Emean) {
  return ssc_overlap(R, IE, Rinit, dB, Emean);
}
double ora_msc_overlap(int w, int rg, const uint8_t* active, const int32_t* IE, int Rinit,
                       int R, double dB, double Emean) {
  return msc_overlap(w, rg, active, IE, Rinit, R, dB, Emean);
}
void ora_ssc_averages(int R, const int32_t* IE, const int32_t* M, double out[5]) {
  ssc_averages(R, IE, M, out);
}
void ora_msc_averages(int w, int rg, const uint8_t* active, const int32_t* IE, const int32_t* M,
                      double out[5]) {
  msc_averages(w, rg, active, IE, M, out);
}
void ora_ssc_resample(int L, int Rold, const int8_t* spins, const uint32_t* Roff,
                      int8_t* spins_new) {
  ssc_resample(Geo(L), Rold, spins, Roff, spins_new);
}
void ora_msc_resample(int L, int w, int rg_old, const void* words, const uint8_t* active,
                      const int32_t* IE, const uint32_t* Roff, int Rnew, void* words_new,
                      uint8_t* active_new, int32_t* IE_new) {
  by_width(w, [&](auto* t) {
    using W = std::remove_pointer_t<decltype(t)>;
    msc_resample<W>(Geo(L), rg_old, (const W*)words, active, IE, Roff, Rnew, (W*)words_new,
                    active_new, IE_new);
  });
}
void ora_ssc_table(double beta, double T[10]) { ssc_table(beta, T); }
void ora_msc_table(double beta, uint32_t T[2]) { msc_table(beta, T); }

}  // extern "C"
