// gfx950 kernels of the population-annealing engine.
//
// One temperature step of one or many runs is the launch sequence
//   k_qgroup (MSC) / k_qpart (SSC) -> k_tau -> k_scan -> k_resample -> k_sweep_{ssc,msc} -> k_avg
// Where a stage needs every block of the previous one (the q sums, the step record), the
// last block of a run to finish does it (last_block), which saves a launch per stage.
// Every kernel maps blockIdx.x to (run, local block) through a prefix table, so the same
// launches serve a single run (plain path) and a batch of independent runs.  All
// population-size-dependent decisions read the device RunState, never the host.
//
// Reference counterparts (PAisingSSC.cu = SSC, PAisingMSC.cu = MSC):
//   k_init_*   ReplicaInit + energyKer       SSC:150-157,195-215 / MSC:169-188,268-317
//   k_qpart / k_qgroup  QKer parts 0 and 1   SSC:217-224 / MSC:319-335
//   k_tau      QKer final part, host lnQ, CalcTauKer   SSC:225-254,533 / MSC:336-360,696
//   k_scan     CalcParSum and the parent of every child  SSC:256-270 / MSC:362-395
//   k_resample resampleKer                   SSC:272-280 / MSC:397-435
//   k_sweep_*  checkKerALL + energyKer       SSC:159-215 / MSC:190-317
//   k_avg      CalcAverages + host thermodynamics bookkeeping  SSC:282-292,614-632

#include <hip/hip_runtime.h>

#include "pa_dev.h"
#include "pa_prims.h"

namespace pa {

constexpr int kRB = kRedBlock;  // 1024

__device__ __forceinline__ int find_run(const int32_t* __restrict__ off, int nruns, int b) {
  int lo = 0, hi = nruns;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int ceil_div(int a, int b) { return (a + b - 1) / b; }

// ---- reference-order FP64 block sum (SSC:107-124) -----------------------------------
// Warps of 32 lanes (two per wave64) reduce by shuffle-down 16,8,4,2,1; lane 0 of each warp
// stores its sum, then warp 0 reduces the warp sums in the same tree.  Result in thread 0.
__device__ __forceinline__ double warp32_tree(double v) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_down(v, off, 32);
  return v;
}

__device__ double ref_block_sum(double v, double* sh) {
  const int lane = threadIdx.x & 31, wid = threadIdx.x >> 5;
  v = warp32_tree(v);
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  const int nw = blockDim.x >> 5;
  v = (threadIdx.x < (unsigned)nw) ? sh[lane] : 0.0;
  if (wid == 0) v = warp32_tree(v);
  return v;
}

// Integer block sums (order-free).
template <typename T>
__device__ T block_sum_int(T v, T* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  return r;
}

// K tree sums at once, each in ref_block_sum's order (sh: 32 K doubles); valid in thread 0.
template <int K>
__device__ void ref_block_sums(double (&v)[K], double* sh) {
  const int lane = threadIdx.x & 31, wid = threadIdx.x >> 5;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = warp32_tree(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k * 32 + wid] = v[k];
  __syncthreads();
  const int nw = blockDim.x >> 5;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = (threadIdx.x < (unsigned)nw) ? sh[k * 32 + lane] : 0.0;
  if (wid == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = warp32_tree(v[k]);
}

// Partials that the last block of the same launch reads (last_block) are written and read
// with device-scope relaxed atomics: these go to the device-coherent level themselves, so
// neither side needs the L2 write-back or invalidation of an agent-scope fence (one per
// block of every run of a batch would cost more than the launch they save).
template <typename T>
__device__ __forceinline__ void put_dev(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T get_dev(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Last-block election.  Every participating block of a run calls this once, with all its
// threads, after writing its partials with put_dev; it returns true in the one block that
// arrives last (whose get_dev reads then see every other block's partials) and re-arms the
// counter for the next launch.  Each thread waits for its own stores to complete before the
// block's arrival is counted.  Blocks that do not participate must decide so from state
// that the last block leaves as it is.
__device__ bool last_block(int32_t* ctr, int nblocks) {
  __shared__ int last_sh;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int prev = __hip_atomic_fetch_add(ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last_sh = prev == nblocks - 1;
    if (prev == nblocks - 1) __hip_atomic_store(ctr, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return last_sh != 0;
}

// ---- geometry --------------------------------------------------------------------------
// Sublattice site idx = ty*L/2 + tx.  Neighbours of a site of sublattice `side` in the other
// sublattice: horizontal i1 (left or right by row parity), the same idx, row below i3 and
// row above i4 (SSC:169-172 for A, SSC:181-184 for B).
struct Site {
  int idx, i1, i3, i4;
};

__device__ __forceinline__ Site site_of(int idx, int side, int ty, int tx, int L, int Ld2,
                                        int Nh) {
  Site s;
  s.idx = idx;
  const int odd = ty & 1;
  const int right = (side == 0) ? odd : !odd;  // A: odd rows look right, B: even rows
  const int x1 = right ? (tx + 1 == Ld2 ? 0 : tx + 1) : (tx == 0 ? Ld2 - 1 : tx - 1);
  s.i1 = idx - tx + x1;
  s.i3 = (ty == 0) ? idx + Nh - Ld2 : idx - Ld2;
  s.i4 = (ty == L - 1) ? idx - Nh + Ld2 : idx + Ld2;
  return s;
}

__device__ __forceinline__ Site site_at(int idx, int side, const RunDesc& d) {
  const int ty = (int)__umulhi((uint32_t)idx, d.ld2_magic);
  const int tx = idx - ty * d.Ld2;
  return site_of(idx, side, ty, tx, d.L, d.Ld2, d.Nh);
}

// ---- Boltzmann acceptance, SSC ----------------------------------------------------------
// u < T  <=>  v < V(T) for the 32-bit word v behind u; V is computed on the host per step
// for T = 1 (dE <= 0), exp(-4 beta), exp(-8 beta).  The kernel keeps the three values as
// 32-bit words in a 9-entry LDS table indexed by prod + 4, prod = s_i * sum of neighbours:
// one LDS read instead of a compare/select chain.  V(1) can be 2^32 (midpoint mapping:
// every u < 1); ALWAYS1 then accepts every dE <= 0 move.
template <bool ALWAYS1>
__device__ __forceinline__ void ssc_flip(int8_t* own, const int8_t* oth, const Site& q,
                                         uint32_t v, const uint32_t* tab) {
  const int sp = own[q.idx];
  const int prod = sp * (oth[q.i1] + oth[q.idx] + oth[q.i3] + oth[q.i4]);
  bool acc = v < tab[prod + 4];
  if (ALWAYS1) acc = acc || prod <= 0;
  own[q.idx] = (int8_t)(acc ? -sp : sp);
}

// Four sites of one sublattice with the four words of one Philox block.  sp holds the
// sites' spins (a site is only ever written by the thread that updates it, so the caller
// may keep them in registers across sweeps).  All LDS reads are issued before the first
// store: the stores go to the other sublattice, which the compiler cannot prove.
template <bool ALWAYS1>
__device__ __forceinline__ void ssc_flip4(int8_t* own, const int8_t* oth, const Site (&q)[4],
                                          int (&sp)[4], const U4& rb, const uint32_t* tab) {
  int nb[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    nb[j] = oth[q[j].i1] + oth[q[j].idx] + oth[q[j].i3] + oth[q[j].i4];
  const uint32_t v[4] = {rb.x, rb.y, rb.z, rb.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int prod = sp[j] * nb[j];
    bool acc = v[j] < tab[prod + 4];
    if (ALWAYS1) acc = acc || prod <= 0;
    sp[j] = acc ? -sp[j] : sp[j];
    own[q[j].idx] = (int8_t)sp[j];
  }
}

// ---- MSC word update (MSC:207-229) -----------------------------------------------------
// Bit-sliced count of antiparallel neighbours, one LCG value per replica bit, acceptance
// mask: count >= 2 always, count 1 iff ran < T4, count 0 iff ran < T8.
template <typename S>
__device__ __forceinline__ S msc_acc(S s, S n1, S n2, S n3, S n4, S c4, S c8) {
  const S I1 = s ^ n1, I2 = s ^ n2, I3 = s ^ n3, I4 = s ^ n4;
  const S x12 = I1 ^ I2, x34 = I3 ^ I4, a12 = I1 & I2, a34 = I3 & I4;
  const S sum0 = x12 ^ x34;
  const S sum1 = (x12 & x34) ^ a12 ^ a34;
  const S sum2 = a12 & a34;
  const S s12 = sum1 | sum2;
  return s12 | (~s12 & ((sum0 & c4) | (~sum0 & c8)));
}

// One LCG step, ran' = A ran + C (mod 2^32), as a single v_mad_u64_u32 (A in an SGPR, C as
// a 64-bit VGPR pair); the compiler would otherwise emit v_mul_lo_u32 + v_add_u32.
__device__ __forceinline__ uint32_t lcg_next(uint32_t r, uint64_t c64) {
  uint64_t p, carry;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(p), "=s"(carry) : "s"(kLcgA), "v"(r), "v"(c64));
  return (uint32_t)p;
}

// a4 = 2 a4 + [r < T4], a8 = 2 a8 + [r < T8]: a compare into an SGPR lane mask and an add
// with carry-in per mask (the compiler's form is compare + select + or).  The s_nop keeps
// two wait states between each mask write and its read as carry-in.
__device__ __forceinline__ void lcg_cmp2(uint32_t& a4, uint32_t& a8, uint32_t r, uint32_t T4,
                                         uint32_t T8) {
  uint64_t m4, m8, c0, c1;
  asm("v_cmp_gt_u32_e64 %[m4], %[t4], %[r]\n\t"
      "v_cmp_gt_u32_e64 %[m8], %[t8], %[r]\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32_e64 %[a4], %[c0], %[a4], %[a4], %[m4]\n\t"
      "v_addc_co_u32_e64 %[a8], %[c1], %[a8], %[a8], %[m8]"
      : [a4] "+v"(a4), [a8] "+v"(a8), [m4] "=&s"(m4), [m8] "=&s"(m8), [c0] "=&s"(c0),
        [c1] "=&s"(c1)
      : [t4] "s"(T4), [t8] "s"(T8), [r] "v"(r));
}

// Four chains at once, ordered so that every mask is written at least four VALU
// instructions before its carry-in read (no s_nop needed) with four mask pairs and one
// shared dummy carry-out; then the four LCG steps.
__device__ __forceinline__ void lcg_cmp_step4(uint32_t (&a4)[4], uint32_t (&a8)[4],
                                              uint32_t (&r)[4], uint32_t T4, uint32_t T8,
                                              uint64_t c64) {
  uint64_t m0, m1, m2, m3, cc, p[4];
  asm("v_cmp_gt_u32_e64 %[m0], %[t4], %[r0]\n\t"
      "v_cmp_gt_u32_e64 %[m1], %[t8], %[r0]\n\t"
      "v_cmp_gt_u32_e64 %[m2], %[t4], %[r1]\n\t"
      "v_cmp_gt_u32_e64 %[m3], %[t8], %[r1]\n\t"
      "v_addc_co_u32_e64 %[a40], %[cc], %[a40], %[a40], %[m0]\n\t"
      "v_addc_co_u32_e64 %[a80], %[cc], %[a80], %[a80], %[m1]\n\t"
      "v_cmp_gt_u32_e64 %[m0], %[t4], %[r2]\n\t"
      "v_cmp_gt_u32_e64 %[m1], %[t8], %[r2]\n\t"
      "v_addc_co_u32_e64 %[a41], %[cc], %[a41], %[a41], %[m2]\n\t"
      "v_addc_co_u32_e64 %[a81], %[cc], %[a81], %[a81], %[m3]\n\t"
      "v_cmp_gt_u32_e64 %[m2], %[t4], %[r3]\n\t"
      "v_cmp_gt_u32_e64 %[m3], %[t8], %[r3]\n\t"
      "v_addc_co_u32_e64 %[a42], %[cc], %[a42], %[a42], %[m0]\n\t"
      "v_addc_co_u32_e64 %[a82], %[cc], %[a82], %[a82], %[m1]\n\t"
      "v_addc_co_u32_e64 %[a43], %[cc], %[a43], %[a43], %[m2]\n\t"
      "v_addc_co_u32_e64 %[a83], %[cc], %[a83], %[a83], %[m3]\n\t"
      "v_mad_u64_u32 %[p0], %[cc], %[A], %[r0], %[C]\n\t"
      "v_mad_u64_u32 %[p1], %[cc], %[A], %[r1], %[C]\n\t"
      "v_mad_u64_u32 %[p2], %[cc], %[A], %[r2], %[C]\n\t"
      "v_mad_u64_u32 %[p3], %[cc], %[A], %[r3], %[C]"
      : [a40] "+v"(a4[0]), [a80] "+v"(a8[0]), [a41] "+v"(a4[1]), [a81] "+v"(a8[1]),
        [a42] "+v"(a4[2]), [a82] "+v"(a8[2]), [a43] "+v"(a4[3]), [a83] "+v"(a8[3]),
        [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3), [cc] "=&s"(cc),
        [p0] "=&v"(p[0]), [p1] "=&v"(p[1]), [p2] "=&v"(p[2]), [p3] "=&v"(p[3])
      : [t4] "s"(T4), [t8] "s"(T8), [r0] "v"(r[0]), [r1] "v"(r[1]), [r2] "v"(r[2]),
        [r3] "v"(r[3]), [A] "s"(kLcgA), [C] "v"(c64));
#pragma unroll
  for (int j = 0; j < 4; ++j) r[j] = (uint32_t)p[j];
}

// Condition masks for SW consecutive LCG values starting at `ran` (NW independent chains
// interleaved for instruction-level parallelism): bit i of c4 (c8) is [ran_i < T4]
// ([ran_i < T8]).  The masks are accumulated most-significant-first and bit-reversed.
template <typename S, int NW>
__device__ __forceinline__ void lcg_masks(uint32_t (&ran)[NW], uint32_t T4, uint32_t T8,
                                          S (&c4)[NW], S (&c8)[NW]) {
  constexpr int SW = (int)sizeof(S) * 8;
  constexpr int HB = SW < 32 ? SW : 32;  // bits per 32-bit accumulator
  uint64_t c64 = kLcgC;
#pragma unroll
  for (int j = 0; j < NW; ++j) { c4[j] = 0; c8[j] = 0; }
#pragma unroll
  for (int half = 0; half < SW / HB; ++half) {
    uint32_t a4[NW], a8[NW];
#pragma unroll
    for (int j = 0; j < NW; ++j) { a4[j] = 0; a8[j] = 0; }
#pragma unroll
    for (int i = 0; i < HB; ++i) {
      if constexpr (NW == 4) {
        lcg_cmp_step4(a4, a8, ran, T4, T8, c64);
      } else {
#pragma unroll
        for (int j = 0; j < NW; ++j) {
          lcg_cmp2(a4[j], a8[j], ran[j], T4, T8);
          ran[j] = lcg_next(ran[j], c64);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      c4[j] |= (S)((uint64_t)(__builtin_bitreverse32(a4[j]) >> (32 - HB)) << (32 * half));
      c8[j] |= (S)((uint64_t)(__builtin_bitreverse32(a8[j]) >> (32 - HB)) << (32 * half));
    }
  }
}

// Condition masks of one word's SW LCG values as four chains of SW/4 steps each, started at
// ran_0, ran_{SW/4}, ran_{SW/2}, ran_{3SW/4} by compile-time jumps (three multiply-adds), so
// that a thread with a single word in flight still has four independent chains.
constexpr uint32_t lcg_jump_a(int k) {
  uint32_t a = 1u;
  for (int i = 0; i < k; ++i) a *= kLcgA;
  return a;
}
constexpr uint32_t lcg_jump_c(int k) {
  uint32_t c = 0u;
  for (int i = 0; i < k; ++i) c = kLcgA * c + kLcgC;
  return c;
}
template <uint32_t A, uint32_t C>
__device__ __forceinline__ uint32_t lcg_apply(uint32_t r) {
  uint64_t p, carry;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(p), "=s"(carry) : "s"(A), "v"(r), "v"((uint64_t)C));
  return (uint32_t)p;
}
template <typename S>
__device__ __forceinline__ void lcg_masks_split(uint32_t r0, uint32_t T4, uint32_t T8, S& c4, S& c8) {
  constexpr int Q = (int)sizeof(S) * 2;  // steps per chain: SW / 4
  static_assert(Q >= 4 && Q <= 16, "16-, 32- or 64-bit slices");
  uint32_t ch[4] = {r0, lcg_apply<lcg_jump_a(Q), lcg_jump_c(Q)>(r0),
                    lcg_apply<lcg_jump_a(2 * Q), lcg_jump_c(2 * Q)>(r0),
                    lcg_apply<lcg_jump_a(3 * Q), lcg_jump_c(3 * Q)>(r0)};
  uint32_t a4[4] = {0u, 0u, 0u, 0u}, a8[4] = {0u, 0u, 0u, 0u};
  const uint64_t c64 = kLcgC;
#pragma unroll
  for (int i = 0; i < Q; ++i) lcg_cmp_step4(a4, a8, ch, T4, T8, c64);
  c4 = 0;
  c8 = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    c4 |= (S)((uint64_t)(__builtin_bitreverse32(a4[j]) >> (32 - Q)) << (Q * j));
    c8 |= (S)((uint64_t)(__builtin_bitreverse32(a8[j]) >> (32 - Q)) << (Q * j));
  }
}

// LCG jump: ran_j = A_j ran_0 + C_j (mod 2^32).
__device__ __forceinline__ void lcg_jump(int j, uint32_t& Aj, uint32_t& Cj) {
  uint32_t a = 1u, c = 0u;
  for (int i = 0; i < j; ++i) { c = kLcgA * c + kLcgC; a = kLcgA * a; }
  Aj = a;
  Cj = c;
}

// ============================================================================================
// Initial population + energies.  mode is implicit: launched once per run before step 1.
// ============================================================================================

// ReplicaInit SSC:150-157 (stream IS0 + t + 128 B; words 2k -> gA, 2k+1 -> gB; spin up iff
// u >= 0.5) followed by energyKer SSC:195-215.
__global__ __launch_bounds__(256) void k_init_ssc(const RunDesc* __restrict__ runs,
                                                  const int32_t* __restrict__ offs, int nruns,
                                                  const uint64_t* __restrict__ is0) {
  const int* off = offs + kKUnits * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  const int B = blockIdx.x - off[r];
  if (B >= d.st->R) return;
  extern __shared__ int8_t lat8[];
  const int Nh = d.Nh, L = d.L, Ld2 = d.Ld2;
  const uint64_t base = is0[r] + (uint64_t)B * kStreamsPerUnit;
  for (int t = threadIdx.x; t < kStreamsPerUnit; t += blockDim.x) {
    U4 blk{0, 0, 0, 0};
    uint64_t cached = ~0ull;
    int k = 0;
    for (int idx = t; idx < Nh; idx += kStreamsPerUnit, ++k) {
      for (int side = 0; side < 2; ++side) {
        const uint64_t p = 2ull * k + side;
        if ((p >> 2) != cached) { cached = p >> 2; blk = stream_block(d.seed, base + t, cached); }
        const uint32_t v = u4_word(blk, (int)(p & 3));
        lat8[side * Nh + idx] = uniform_double(v, d.umap) >= 0.5 ? 1 : -1;
      }
    }
  }
  __syncthreads();
  int e = 0, m = 0;
  for (int idx = threadIdx.x; idx < Nh; idx += blockDim.x) {
    const int ty = idx / Ld2, tx = idx - ty * Ld2;
    const int iL = ty * Ld2 + (tx == 0 ? Ld2 - 1 : tx - 1);
    const int iU = (ty == L - 1 ? 0 : ty + 1) * Ld2 + tx;
    const int sA = lat8[idx], sB = lat8[Nh + idx];
    int Ai2, Bi2;
    if (ty & 1) { Ai2 = sB; Bi2 = lat8[iL]; } else { Ai2 = lat8[Nh + iL]; Bi2 = sA; }
    const int Ai4 = lat8[Nh + iU], Bi4 = lat8[iU];
    e += -sA * (Ai2 + Ai4) - sB * (Bi2 + Bi4);
    m += sA + sB;
  }
  __shared__ int shi[16];
  const int E = block_sum_int<int>(e, shi);
  const int Mg = block_sum_int<int>(m, shi);
  if (threadIdx.x == 0) { d.IE[B] = E; d.M[B] = Mg; }
  int8_t* g = (int8_t*)d.pop[0] + (size_t)B * d.N;
  for (int i = threadIdx.x; i < d.N; i += blockDim.x) g[i] = lat8[i];
}

// Per-lane energy and magnetisation of an MSC lattice slice held in LDS (MSC:268-317):
// per site pair e = 2 (#antiparallel of 4 bonds) - 4, m = 2 (bA + bB) - 2, summed per lane.
template <typename S>
__device__ void msc_energy_slice(const S* lat, int L, int Ld2, int Nh, int* accE, int* accM,
                                 int lane0, int32_t* IE, int32_t* Mo, size_t lane_base) {
  constexpr int SW = (int)sizeof(S) * 8;
  for (int i = threadIdx.x; i < SW; i += blockDim.x) { accE[i] = 0; accM[i] = 0; }
  __syncthreads();
  const int nq = blockDim.x / SW > 0 ? blockDim.x / SW : 1;
  const int lane = threadIdx.x % SW, q = threadIdx.x / SW;
  if (q < nq) {
    int e = 0, m = 0;
    for (int idx = q; idx < Nh; idx += nq) {
      const int ty = idx / Ld2, tx = idx - ty * Ld2;
      const int iL = ty * Ld2 + (tx == 0 ? Ld2 - 1 : tx - 1);
      const int iU = (ty == L - 1 ? 0 : ty + 1) * Ld2 + tx;
      const S sA = lat[idx], sB = lat[Nh + idx];
      S Ai2, Bi2;
      if (ty & 1) { Ai2 = sB; Bi2 = lat[iL]; } else { Ai2 = lat[Nh + iL]; Bi2 = sA; }
      const S Ai4 = lat[Nh + iU], Bi4 = lat[iU];
      const S I1 = sA ^ Ai2, I2 = sA ^ Ai4, I3 = sB ^ Bi2, I4 = sB ^ Bi4;
      const int c = (int)((I1 >> lane) & 1) + (int)((I2 >> lane) & 1) + (int)((I3 >> lane) & 1) +
                    (int)((I4 >> lane) & 1);
      e += 2 * c - 4;
      m += 2 * ((int)((sA >> lane) & 1) + (int)((sB >> lane) & 1)) - 2;
    }
    atomicAdd(&accE[lane], e);
    atomicAdd(&accM[lane], m);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SW; i += blockDim.x) {
    IE[lane_base + lane0 + i] = accE[i];
    Mo[lane_base + lane0 + i] = accM[i];
  }
}

// ReplicaInit MSC:169-188 (w<=32: words 2k, 2k+1, of which w = 8 / 16 keep the low byte /
// half, MSC:173-178; w=64: (word 4k << 32) + word 4k+1 and (word 4k+2 << 32) + word 4k+3,
// first draw in the high half), written straight to HBM; the energies follow from a
// zero-sweep pass of k_sweep_msc.
template <typename W>
__global__ __launch_bounds__(128) void k_init_msc(const RunDesc* __restrict__ runs,
                                                  const int32_t* __restrict__ offs, int nruns,
                                                  const uint64_t* __restrict__ is0) {
  constexpr int w = (int)sizeof(W) * 8;
  const int* off = offs + kKUnits * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  const int B = blockIdx.x - off[r];
  const int R0 = d.st->R;
  if (B >= ceil_div(R0, w)) return;
  const int Nh = d.Nh;
  W* gl = (W*)d.pop[0] + (size_t)B * d.N;
  const uint64_t base = is0[r] + (uint64_t)B * kStreamsPerUnit;
  constexpr int per = (w <= 32) ? 1 : 2;  // stream words per lattice word
  const int t = threadIdx.x;
  U4 blk{0, 0, 0, 0};
  uint64_t cached = ~0ull;
  int k = 0;
  for (int idx = t; idx < Nh; idx += kStreamsPerUnit, ++k) {
    for (int side = 0; side < 2; ++side) {
      W val = 0;
      for (int h = 0; h < per; ++h) {
        const uint64_t p = (uint64_t)(2 * k + side) * per + h;
        if ((p >> 2) != cached) { cached = p >> 2; blk = stream_block(d.seed, base + t, cached); }
        val = (per == 1) ? (W)u4_word(blk, (int)(p & 3))
                         : (W)(((uint64_t)val << 32) + u4_word(blk, (int)(p & 3)));
      }
      gl[side * Nh + idx] = val;
    }
  }
  if (t == 0) {
    uint64_t mask = 0;
    for (int i = 0; i < w; ++i)
      if ((int64_t)B * w + i < R0) mask |= 1ull << i;
    d.amask[0][B] = mask;
  }
}

// ============================================================================================
// Q partial sums: exp(-dB (E - Emean)) summed in the reference tree order.
// MSC lane sums (QKer part 0, MSC:321-329): a block of w threads per group; here one lane
// per thread, the warp-32 shuffle tree per 32 lanes, and for w = 64 the two warp sums
// added as warp 0 of a 64-thread block does.  For w = 8 and 16 the reference reduces with
// smallblockReduceSum (MSC:96-105): shared[t] += shared[t + stride], stride = w/2 .. 1,
// which a shuffle-down tree of width w reproduces in lane 0 of each group.
// ============================================================================================
__global__ __launch_bounds__(1024) void k_qgroup(const RunDesc* __restrict__ runs,
                                                 const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKChunks * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (d.coding != kMSC || !st->active || st->overflow) return;
  const int c = blockIdx.x - off[r];
  const int R = st->R, w = d.w, rg = ceil_div(R, w);
  if (c >= ceil_div(rg * w, kRB)) return;
  const int n = c * kRB + threadIdx.x;
  const int g = n / w, lane = n % w;
  const double dB = d.dbeta[st->step], Emean = st->Emean;
  double f = 0.0;
  if (g < rg && ((d.amask[st->buf][g] >> lane) & 1))
    f = exp_fp64(-dB * ((double)d.IE[n] - Emean));
  if (w < 32) {
    for (int o = w >> 1; o > 0; o >>= 1) f += __shfl_down(f, o, w);
  } else {
    f = warp32_tree(f);
    if (w == 64) {
      const double hi = __shfl_down(f, 32, 64);
      f = f + hi;
    }
  }
  if (lane == 0 && g < rg) put_dev(d.gsum + g, f);
  // QKer part 1 (MSC:330-335) in the run's last block: tree sums of 1024 group sums each
  if (!last_block(&st->qdone, ceil_div(rg * w, kRB))) return;
  __shared__ double sh[32];
  for (int b = 0; b * kRB < rg; ++b) {
    const int gi = b * kRB + threadIdx.x;
    const double s = ref_block_sum(gi < rg ? get_dev(d.gsum + gi) : 0.0, sh);
    if (threadIdx.x == 0) d.qblk[b] = s;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    StepOut& o = d.out[st->step];
    o.roff_hash = 0;
    o.spin_hash = 0;
  }
}

// SSC (QKer part 0, SSC:220-224): one block per 1024 replicas (MSC populations: k_qgroup).
__global__ __launch_bounds__(1024) void k_qpart(const RunDesc* __restrict__ runs,
                                                const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKGroupChunks * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (d.coding != kSSC || !st->active || st->overflow) return;
  const int c = blockIdx.x - off[r];
  const int R = st->R, step = st->step;
  const double dB = d.dbeta[step], Emean = st->Emean;
  __shared__ double sh[32];
  if (c >= ceil_div(R, kRB)) return;
  const int idx = c * kRB + threadIdx.x;
  const double f = (idx < R) ? exp_fp64(-dB * ((double)d.IE[idx] - Emean)) : 0.0;
  const double s = ref_block_sum(f, sh);
  if (threadIdx.x == 0) {
    d.qblk[c] = s;
    if (c == 0) {
      StepOut& o = d.out[step];
      o.roff_hash = 0;
      o.spin_hash = 0;
    }
  }
}

// ============================================================================================
// k_tau: q from the block partials (QKer last part, SSC:225-233 / MSC:336-344), lnQ
// (SSC:533), then copy counts (CalcTauKer, SSC:236-254 / MSC:347-360):
//   mu = Rinit/R exp(-dB E - lnQ), Roff = floor(mu) + [u < mu - floor(mu)],
//   u from word 0 of subsequence IS + n (n = replica, or group*w + lane for MSC).
// One block per 1024 population members; every block evaluates q itself (identical order).
// ============================================================================================
__global__ __launch_bounds__(1024) void k_tau(const RunDesc* __restrict__ runs,
                                              const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKChunks * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  const int c = blockIdx.x - off[r];
  const int R = st->R, step = st->step;
  const int w = d.w;
  const int nitems = (d.coding == kSSC) ? R : ceil_div(R, w);  // QKer's reduction items
  const int nel = (d.coding == kSSC) ? R : ceil_div(R, w) * w;  // copy-count slots
  if (c >= ceil_div(nel, kRB)) return;
  __shared__ double sh[32];
  __shared__ uint32_t shu[16];
  __shared__ uint64_t shh[16];
  const int nqb = ceil_div(nitems, kRB);
  double acc = 0.0;
  for (int j = 0; j * kRB < nitems; j += kRB) {
    const int b = threadIdx.x + j;
    const double f = (b < nqb) ? d.qblk[b] : 0.0;
    acc += ref_block_sum(f, sh);
    __syncthreads();
  }
  __shared__ double lnq_sh;
  if (threadIdx.x == 0) {
    const double dB = d.dbeta[step];
    const double lnQ = -(dB)*st->Emean + log_fp64(acc) - log_fp64((double)R);
    lnq_sh = lnQ;
    if (c == 0) {
      d.out[step].q = acc;
      d.out[step].lnQ = lnQ;
    }
  }
  __syncthreads();
  const double lnQ = lnq_sh;
  const double dB = d.dbeta[step];
  const double ratio = (double)d.Rinit / (double)R;
  const int n = c * kRB + threadIdx.x;
  uint32_t roff = 0;
  uint64_t hv = 0;
  if (n < nel) {
    bool act = true;
    if (d.coding == kMSC) act = (d.amask[st->buf][n / w] >> (n % w)) & 1;
    if (act) {
      const double mu = ratio * exp_fp64(fma_(-dB, (double)d.IE[n], -lnQ));
      const double mf = floor(mu);
      const uint32_t v = stream_block(d.seed, st->IS + (uint64_t)n, 0).x;
      roff = (uniform_double(v, d.umap) < (mu - mf)) ? (uint32_t)(mf + 1.0) : (uint32_t)mf;
    }
    d.Roff[n] = roff;
    if (d.trace) hv = mix64(mix64((uint64_t)n) ^ (uint64_t)roff);
  }
  const uint32_t tot = block_sum_int<uint32_t>(roff, shu);
  if (threadIdx.x == 0) d.cblk[c] = tot;
  if (d.trace) {
    const uint64_t ht = block_sum_int<uint64_t>(hv, shh);
    if (threadIdx.x == 0) atomicAdd((unsigned long long*)&d.out[step].roff_hash, (unsigned long long)ht);
  }
}

// ============================================================================================
// k_overlap: energy-histogram overlap of the current population for the candidate step
// dB = dbeta[step] (adaptive step; HistogramOverlap + CalcOverlap, SSC:308-334 and
// MSC:466-499): q from the k_qgroup/k_qpart partials as in k_tau, lnQ, then per block of
// 1024 replicas (SSC) or (1024 groups, lane) (MSC) the tree sum of
// min(1, Rinit/R exp(-dB E - lnQ)) into avblk[5 c].  The host adds the block sums in launch
// order (the reference adds them with atomicAdd).
// ============================================================================================
__global__ __launch_bounds__(1024) void k_overlap(const RunDesc* __restrict__ runs,
                                                  const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKAvg * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (st->overflow) return;
  const int c = blockIdx.x - off[r];
  const int R = st->R, w = d.w, buf = st->buf, step = st->step;
  int nb, b, by;
  if (d.coding == kSSC) {
    nb = ceil_div(R, kRB);
    if (c >= nb) return;
    b = c; by = 0;
  } else {
    nb = ceil_div(ceil_div(R, w), kRB);
    if (c >= nb * w) return;
    by = c / nb; b = c % nb;
  }
  __shared__ double sh[32];
  const int nitems = (d.coding == kSSC) ? R : ceil_div(R, w);
  const int nqb = ceil_div(nitems, kRB);
  double acc = 0.0;
  for (int j = 0; j * kRB < nitems; j += kRB) {
    const int k = threadIdx.x + j;
    const double f = (k < nqb) ? d.qblk[k] : 0.0;
    acc += ref_block_sum(f, sh);
    __syncthreads();
  }
  // the block sum is valid in thread 0 only: broadcast lnQ
  const double dB = d.dbeta[step];
  __shared__ double lnq_sh;
  if (threadIdx.x == 0) lnq_sh = -(dB)*st->Emean + log_fp64(acc) - log_fp64((double)R);
  __syncthreads();
  const double lnQ = lnq_sh;
  const double ratio = (double)d.Rinit / (double)R;
  const int idx = b * kRB + threadIdx.x;
  double v = 0.0;
  if (d.coding == kSSC) {
    if (idx < R) v = ratio * exp_fp64(fma_(-dB, (double)d.IE[idx], -lnQ));
  } else if (idx < ceil_div(R, w) && ((d.amask[buf][idx] >> by) & 1)) {
    v = ratio * exp_fp64(fma_(-dB, (double)d.IE[(size_t)idx * w + by], -lnQ));
  }
  v = v < 1.0 ? v : 1.0;
  const double sblk = ref_block_sum(v, sh);
  if (threadIdx.x == 0) d.avblk[(size_t)c * 5] = sblk;
}

// ============================================================================================
// k_hist: adds the population's energies to the run's MHR histogram, bin (2N + E) / 4
// (UpdateShistE, SSC:296-302 / MSC:454-460).  init = 1: the initial population (buffer
// st->buf, R members); otherwise the population just equilibrated by this step (buffer
// st->buf ^ 1, Rnew members; launched before k_avg, whose last block ends the step).  Integer atomics: order-free.
// ============================================================================================
__global__ __launch_bounds__(1024) void k_hist(const RunDesc* __restrict__ runs,
                                               const int32_t* __restrict__ offs, int nruns,
                                               int init) {
  const int* off = offs + kKChunks * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!d.hist || st->overflow || (!init && !st->active)) return;
  const int c = blockIdx.x - off[r];
  const int R = init ? st->R : st->Rnew, buf = init ? st->buf : st->buf ^ 1, w = d.w;
  const int n = c * kRB + threadIdx.x;
  bool in;
  if (d.coding == kSSC) in = n < R;
  else in = n / w < ceil_div(R, w) && ((d.amask[buf][n / w] >> (n % w)) & 1);
  if (in) atomicAdd(&d.hist[(2 * d.N + d.IE[n]) / 4], 1);
}

// ============================================================================================
// k_scan: exclusive prefix of Roff in population order (CalcParSum), the parent of every
// child (children of member p are pre[p] .. pre[p] + Roff[p] - 1, SSC:256-280 /
// MSC:362-435), R_i, capacity check.
// ============================================================================================
__global__ __launch_bounds__(1024) void k_scan(const RunDesc* __restrict__ runs,
                                               const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKChunks * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  const int c = blockIdx.x - off[r];
  const int R = st->R, w = d.w;
  const int nel = (d.coding == kSSC) ? R : ceil_div(R, w) * w;
  const int nch = ceil_div(nel, kRB);
  if (c >= nch) return;
  __shared__ uint32_t shu[16];
  __shared__ uint32_t wsum[16];
  // offset of this chunk: sum of the preceding chunk totals
  uint32_t part = 0;
  for (int b = threadIdx.x; b < c; b += blockDim.x) part += d.cblk[b];
  const uint32_t base = block_sum_int<uint32_t>(part, shu);
  __shared__ uint32_t base_sh;
  if (threadIdx.x == 0) base_sh = base;
  __syncthreads();
  const int n = c * kRB + threadIdx.x;
  const uint32_t v = (n < nel) ? d.Roff[n] : 0u;
  // inclusive wave scan
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  uint32_t wpre = 0;
  for (int i = 0; i < wv; ++i) wpre += wsum[i];
  const uint32_t excl = base_sh + wpre + x - v;
  if (n < nel) {
    d.pre[n] = excl;
    // children excl .. excl + v - 1 of member n, in parent order (resampleKer's parent
    // search); past the capacity the run overflows and stops, so those are not written
    const uint32_t cap = (uint32_t)d.cap_units * (uint32_t)(d.coding == kSSC ? 1 : w);
    for (uint32_t i = 0; i < v && excl + i < cap; ++i) d.par[excl + i] = n;
  }
  if (c == nch - 1 && threadIdx.x == blockDim.x - 1) {
    const uint32_t Rn = excl + v;
    st->Rnew = (int)Rn;
    const int units_needed = (d.coding == kSSC) ? (int)Rn : ceil_div((int)Rn, w);
    if (Rn == 0 || units_needed > d.cap_units) st->overflow = 1;
  }
}

// SSC (resampleKer SSC:272-280): child j receives its parent's N spins; all N sites are
// copied (the reference grid N/2/128 drops N/2 mod 128 sites when L is not a multiple
// of 16).  One block per child.
__global__ __launch_bounds__(256) void k_resample(const RunDesc* __restrict__ runs,
                                                  const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKResample * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (d.coding != kSSC || !st->active || st->overflow) return;
  const int j = blockIdx.x - off[r];
  if (j >= st->Rnew) return;
  const int p = d.par[j];
  const int buf = st->buf;
  const uint8_t* src = (const uint8_t*)d.pop[buf] + (size_t)p * d.N;
  uint8_t* dst = (uint8_t*)d.pop[buf ^ 1] + (size_t)j * d.N;
  if ((d.N & 15) == 0) {
    const uint4* s4 = (const uint4*)src;
    uint4* d4 = (uint4*)dst;
    for (int i = threadIdx.x; i < d.N / 16; i += blockDim.x) d4[i] = s4[i];
  } else {
    for (int i = threadIdx.x; i < d.N; i += blockDim.x) dst[i] = src[i];
  }
}

// MSC (resampleKer MSC:397-435): child n = prefix + p lives in new group n mod rg_new, lane
// n div rg_new.  One block per chunk of S site words (S * sizeof(W) = 64 bytes where the
// parents' chunks fit in LDS): the chunk of every parent group is staged in LDS once, then
// every new word of the chunk gathers its w bits from there.  Each byte of the parent
// population is read once and each byte of the children written once.
// NT threads per workgroup, about one new group per thread: 1024 for populations of 32-bit
// words with more than 256 groups (their large parent chunk leaves room for one workgroup
// per CU, which 256 threads cannot keep busy), 256 otherwise (small populations; 64-bit
// words, whose 64 parent indices per thread need the registers).  Any NT is correct: the
// new groups of a chunk are covered by a strided loop.
template <typename W, int NT>
__global__ __launch_bounds__(NT) void k_resample_msc(const RunDesc* __restrict__ runs,
                                                      const int32_t* __restrict__ offs, int nruns) {
  constexpr int w = (int)sizeof(W) * 8;
  const int* off = offs + kKResample * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  if (d.coding != kMSC || d.w != w) return;
  // The run's blocks come in sets of 8 * rs_lp * rs_split starting at a multiple of 8.  The
  // blocks of the rs_lp chunks of one 128-byte line (and their rs_split splits) are 8 apart,
  // so they share an XCD and the line is fetched into that XCD's L2 once.
  const int cb = blockIdx.x - off[r];
  const int per = 8 * d.rs_lp * d.rs_split;
  const int q = cb / per, rem = cb - q * per, t = rem >> 3;
  const int c = (q * 8 + (rem & 7)) * d.rs_lp + t % d.rs_lp, js = t / d.rs_lp;
  const int S = d.rs_sites, N = d.N;
  const int s0 = c * S;
  if (s0 >= N) return;
  const int ns = min(S, N - s0);
  const int R = st->R, Rn = st->Rnew, buf = st->buf;
  const int rgo = ceil_div(R, w), rgn = ceil_div(Rn, w);
  extern __shared__ __align__(16) uint8_t smem_rs[];
  W* old = (W*)smem_rs;
  const W* src = (const W*)d.pop[buf];
  W* dst = (W*)d.pop[buf ^ 1];
  constexpr int SM = 64 / (int)sizeof(W);  // largest chunk: 64 bytes per group
  // whole 64-byte chunks (the common case) take a branch-free gather with vector LDS reads;
  // lanes beyond R_new read a zero chunk staged after the parents'
  const bool whole = S == SM && ns == SM;
  for (int i = threadIdx.x; i < rgo * S; i += blockDim.x) {
    const int g = i / S, s = i - g * S;
    if (s < ns) old[i] = src[(size_t)g * N + s0 + s];
  }
  if (whole && threadIdx.x < SM) old[rgo * SM + threadIdx.x] = 0;
  __syncthreads();
  for (int j = js * blockDim.x + threadIdx.x; j < rgn; j += d.rs_split * blockDim.x) {
    // parents of the w lanes of new group j: independent loads issued together
    int pl[w];
#pragma unroll
    for (int k = 0; k < w; ++k) {
      const int64_t n = (int64_t)k * rgn + j;
      pl[k] = (n < Rn) ? d.par[n] : -1;
    }
    if (whole) {
      W acc[SM];
#pragma unroll
      for (int s = 0; s < SM; ++s) acc[s] = 0;
#pragma unroll
      for (int k = 0; k < w; ++k) {
        const int p = pl[k];
        const unsigned pg = p < 0 ? (unsigned)rgo : (unsigned)p / w;
        const unsigned pb = p < 0 ? 0u : (unsigned)p % w;
        const W* o = old + pg * SM;
#pragma unroll
        for (int s = 0; s < SM; ++s) acc[s] |= ((o[s] >> pb) & (W)1) << k;
      }
      W* out = dst + (size_t)j * N + s0;
#pragma unroll
      for (int s = 0; s < SM; ++s) out[s] = acc[s];
      if (c == 0) {
        uint64_t m = 0;
        for (int k = 0; k < w; ++k)
          if ((int64_t)k * rgn + j < Rn) m |= 1ull << k;
        d.amask[buf ^ 1][j] = m;
      }
      continue;
    }
    W acc[SM];
#pragma unroll
    for (int s = 0; s < SM; ++s) acc[s] = 0;
#pragma unroll
    for (int k = 0; k < w; ++k) {
      const int p = pl[k];
      if (p < 0) continue;
      const int pg = p / w, pb = p - pg * w;
      const W* o = old + pg * S;
#pragma unroll
      for (int s = 0; s < SM; ++s)
        if (s < ns) acc[s] |= ((o[s] >> pb) & (W)1) << k;
    }
    W* out = dst + (size_t)j * N + s0;
#pragma unroll
    for (int s = 0; s < SM; ++s)
      if (s < ns) out[s] = acc[s];
    if (c == 0) {
      uint64_t m = 0;
      for (int k = 0; k < w; ++k)
        if ((int64_t)k * rgn + j < Rn) m |= 1ull << k;
      d.amask[buf ^ 1][j] = m;
    }
  }
}

// MSC resampling of large populations (resampleKer MSC:397-435 again), by windows of new
// groups.  Children n = k rg_new + j of one lane k and consecutive new groups j are
// consecutive, and children are numbered in parent order, so for each lane the parents of a
// window of J new groups lie in one run of parent groups: about J + 2w groups for all lanes,
// against every parent group for k_resample_msc.  A workgroup (one thread per new group)
// finds those runs from the window's first and last child per lane, keeps the (slot, bit) of
// each of its 32/64 children in registers, and walks a range of whole 128-byte site chunks:
// stage the window's parent rows of the chunk in LDS (slot T: zeros, for lanes beyond
// R_new), gather, write one 128-byte line per new group.  A window whose parents exceed
// kWinSlots rows is processed in halves (a single new group needs at most w rows).  Run
// blocks: windows (capacity / J) x site-chunk ranges (rs_split).
// 32-bit word m (little-endian) of an array of words of type W.
template <typename W, int NW>
__device__ __forceinline__ uint32_t word32_of(const W (&a)[NW], int m) {
  if constexpr (sizeof(W) == 8) return (m & 1) ? (uint32_t)(a[m >> 1] >> 32) : (uint32_t)a[m >> 1];
  else if constexpr (sizeof(W) == 4) return a[m];
  else if constexpr (sizeof(W) == 2) return (uint32_t)a[2 * m] | ((uint32_t)a[2 * m + 1] << 16);
  else
    return (uint32_t)a[4 * m] | ((uint32_t)a[4 * m + 1] << 8) | ((uint32_t)a[4 * m + 2] << 16) |
           ((uint32_t)a[4 * m + 3] << 24);
}

template <typename W>
__global__ __launch_bounds__(256) void k_resample_win(const RunDesc* __restrict__ runs,
                                                      const int32_t* __restrict__ offs, int nruns) {
  constexpr int w = (int)sizeof(W) * 8;
  constexpr int S = 128 / (int)sizeof(W);
  constexpr int J = kWinJ;
  const int* off = offs + kKResample * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  if (d.coding != kMSC || d.w != w) return;
  const int Rn = st->Rnew, buf = st->buf;
  const int rgn = ceil_div(Rn, w);
  const int maxwin = ceil_div(d.cap_units, J);
  const int b = blockIdx.x - off[r];
  const int win = b % maxwin, range = b / maxwin;
  if (win * J >= rgn) return;
  const int N = d.N, nch = N / S;
  const int cpr = ceil_div(nch, d.rs_split);
  const int c0 = range * cpr, c1 = min(nch, c0 + cpr);
  if (c0 >= c1) return;
  extern __shared__ __align__(16) uint8_t smem_rw[];
  uint4* old4 = (uint4*)smem_rw;
  const W* old = (const W*)smem_rw;
  int* slotg = (int*)(smem_rw + (size_t)(kWinSlots + 1) * 128);
  __shared__ int kof[w + 1], kgl[w];
  const W* src = (const W*)d.pop[buf];
  W* dst = (W*)d.pop[buf ^ 1];
  const int jend = min(rgn, (win + 1) * J);
  int cur = win * J;
  while (cur < jend) {
    int jc = jend - cur;
    for (;;) {
      // rows of lane k: parent groups of its first and last child in [cur, cur + jc)
      if (threadIdx.x < w) {
        const int k = threadIdx.x;
        const int64_t nlo = (int64_t)k * rgn + cur;
        const int64_t nhi = min((int64_t)k * rgn + cur + jc, (int64_t)Rn) - 1;
        int glo = 0, cnt = 0;
        if (nlo <= nhi) {
          glo = d.par[nlo] / w;
          cnt = d.par[nhi] / w - glo + 1;
        }
        kgl[k] = glo;
        kof[k + 1] = cnt;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        kof[0] = 0;
        for (int k = 0; k < w; ++k) kof[k + 1] += kof[k];
      }
      __syncthreads();
      if (kof[w] <= kWinSlots || jc == 1) break;
      jc = (jc + 1) / 2;
      __syncthreads();
    }
    const int T = kof[w];
    for (int i = threadIdx.x; i < T; i += blockDim.x) {
      int lo = 0, hi = w - 1;  // last lane k with kof[k] <= i
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (kof[mid] <= i) lo = mid;
        else hi = mid - 1;
      }
      slotg[i] = kgl[lo] + i - kof[lo];
    }
    const int j = cur + (int)threadIdx.x;
    const bool jv = (int)threadIdx.x < jc;
    uint32_t dsc[w / 2];  // (slot << 6 | bit) of lanes 2i and 2i + 1 in the halves of dsc[i]
#pragma unroll
    for (int k = 0; k < w; ++k) {
      const int64_t n = (int64_t)k * rgn + j;
      uint32_t v = (uint32_t)T << 6;
      if (jv && n < Rn) {
        const int p = d.par[n];
        v = ((uint32_t)(kof[k] + p / w - kgl[k]) << 6) | (uint32_t)(p % w);
      }
      if (k & 1) dsc[k >> 1] |= v << 16;
      else dsc[k >> 1] = v;
    }
    if (jv && c0 == 0) {
      uint64_t m = 0;
      for (int k = 0; k < w; ++k)
        if ((int64_t)k * rgn + j < Rn) m |= 1ull << k;
      d.amask[buf ^ 1][j] = m;
    }
    __syncthreads();
    for (int c = c0; c < c1; ++c) {
      const int s0 = c * S;
      for (int idx = threadIdx.x; idx < T * 8; idx += blockDim.x) {
        const int i = idx >> 3, q = idx & 7;
        old4[idx] = ((const uint4*)(src + (size_t)slotg[i] * N + s0))[q];
      }
      if (threadIdx.x < 8) old4[T * 8 + threadIdx.x] = make_uint4(0u, 0u, 0u, 0u);
      __syncthreads();
      if (jv) {
        W acc[S];
#pragma unroll
        for (int s = 0; s < S; ++s) acc[s] = 0;
#pragma unroll
        for (int k = 0; k < w; ++k) {
          const uint32_t v = (dsc[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
          const W* o = old + (v >> 6) * S;
          const unsigned pb = v & 63u;
#pragma unroll
          for (int s = 0; s < S; ++s) acc[s] |= ((o[s] >> pb) & (W)1) << k;
        }
        uint4* out = (uint4*)(dst + (size_t)j * N + s0);
#pragma unroll
        for (int q = 0; q < 8; ++q)
          out[q] = make_uint4(word32_of(acc, 4 * q), word32_of(acc, 4 * q + 1),
                              word32_of(acc, 4 * q + 2), word32_of(acc, 4 * q + 3));
      }
      __syncthreads();
    }
    cur += jc;
  }
}

// ============================================================================================
// Sweeps.  The unit's lattice lives in LDS for all EqSw sweeps of the step.  Site (side h,
// idx) of sweep s is driven by stream t = idx mod 128 of the unit (subsequence
// IS_sweep + t + 128*unit), word p = (2s + h) n_t + idx div 128 with n_t = ceil((N/2-t)/128)
// (SSC:164-188, MSC:195-261).  When N/2 is a multiple of 512 every stream owns whole Philox
// blocks per half sweep, and each thread evaluates one block for four sites at once.
// ============================================================================================

// Stream walk for lattices with at most K sites per stream and half sweep (N/2 <= 128 K):
// the sites' offsets stay in registers for the whole step, packed two 16-bit offsets per
// register, and the stream's words are consumed in order, word p = (2s + h) n_t + k, so a
// running counter replaces the index arithmetic and a Philox block is evaluated every
// fourth word.
template <int K, bool ALWAYS1>
__device__ __forceinline__ void ssc_walk_cached(int8_t* lat8, const RunDesc& d, uint64_t seed,
                                                uint64_t sub0, uint32_t nsw, const uint32_t* tab,
                                                int t, int nt) {
  const int Nh = d.Nh;
  uint32_t pa[2][K], pb[2][K];  // idx | i1 << 16, i3 | i4 << 16
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Site q = site_at(k < nt ? t + kStreamsPerUnit * k : 0, h, d);
      pa[h][k] = (uint32_t)q.idx | ((uint32_t)q.i1 << 16);
      pb[h][k] = (uint32_t)q.i3 | ((uint32_t)q.i4 << 16);
    }
  U4 rb{0, 0, 0, 0};
  uint64_t p = 0;
  for (uint32_t s = 0; s < nsw; ++s) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int8_t* own = lat8 + h * Nh;
      const int8_t* oth = lat8 + (h ^ 1) * Nh;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (k < nt) {
          if ((p & 3) == 0) rb = stream_block(seed, sub0 + t, p >> 2);
          Site q;
          q.idx = (int)(pa[h][k] & 0xFFFFu);
          q.i1 = (int)(pa[h][k] >> 16);
          q.i3 = (int)(pb[h][k] & 0xFFFFu);
          q.i4 = (int)(pb[h][k] >> 16);
          ssc_flip<ALWAYS1>(own, oth, q, u4_word(rb, (int)(p & 3)), tab);
          ++p;
        }
      }
      __syncthreads();
    }
  }
}

// SSC sweeps of one replica held in LDS (lat8: sublattice A then B).  PATH selects the code
// compiled into the kernel (chosen per run on the host by ssc_path), so that the register-
// caching paths do not raise the register count (and lower the occupancy) of the fast path:
//   0: N/2 a multiple of 512 with at most NT/128 Philox blocks per thread and half sweep
//      (fast), or the uncached block loop / stream walk for anything larger;
//   1: aligned, two or three blocks per thread (register-cached block loop);
//   2, 3: not aligned, at most 4 / 9 sites per stream (register-cached stream walk);
//   4: aligned, more blocks per thread, L/2 dividing 128 (each thread on one column);
//   5: not aligned, more than 9 sites per stream (Philox blocks dealt to P threads);
//   6: aligned, columns of a stream's sites repeating every third site (L = 96, 192).
template <int NT, bool ALWAYS1, int PATH>
__device__ __forceinline__ void ssc_sweeps(int8_t* lat8, const RunDesc& d, uint64_t seed,
                                           uint64_t sub0, uint32_t nsw,
                                           const uint32_t* tab) {
  const int Nh = d.Nh;
  if constexpr (PATH == 2 || PATH == 3) {
    const int t = threadIdx.x;
    const int nt = (t < kStreamsPerUnit && t < Nh) ? (Nh - t + kStreamsPerUnit - 1) / kStreamsPerUnit : 0;
    ssc_walk_cached<PATH == 2 ? 4 : 9, ALWAYS1>(lat8, d, seed, sub0, nsw, tab, t, nt);
    return;
  }
  constexpr int P = NT / kStreamsPerUnit;
  // part is uniform per wave (a stream part is two waves): as a scalar, block indices,
  // Philox counters and per-block tests stay on the scalar unit (not in the register-cached
  // block loop, whose offsets are set up per thread and which then needs more registers)
  const int t = threadIdx.x % kStreamsPerUnit;
  const int part = PATH == 1 ? (int)(threadIdx.x / kStreamsPerUnit)
                             : __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kStreamsPerUnit));
  if constexpr (PATH == 4) {
    // N/2 a multiple of 512 with more than 3 blocks per thread and L/2 dividing 128
    // (L = 128, 256): every site of thread t lies in column t % Ld2 and site t + 128 k in row
    // t / Ld2 + k q128, so the horizontal neighbour's offset is fixed per (half, jj) and the
    // vertical neighbours wrap only above jj = 0 and below jj = 3.
    const int Ld2 = d.Ld2, nb = Nh / 512;
    const int tx = t % Ld2, ty0 = t / Ld2;
    const int dR = tx + 1 == Ld2 ? 1 - Ld2 : 1, dL = tx == 0 ? Ld2 - 1 : -1;
    int d1[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) d1[h][jj] = ((ty0 + jj * d.q128 + h) & 1) ? dR : dL;
    for (uint32_t s = 0; s < nsw; ++s) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int8_t* own = lat8 + h * Nh;
        const int8_t* oth = lat8 + (h ^ 1) * Nh;
        for (int m = part; m < nb; m += P) {
          const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m);
          const int o0 = t + 512 * m;
          Site q[4];
          int sp[4];
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            const int o = o0 + jj * kStreamsPerUnit;
            q[jj].idx = o;
            q[jj].i1 = o + d1[h][jj];
            q[jj].i3 = (jj == 0 && o < Ld2) ? o - Ld2 + Nh : o - Ld2;
            q[jj].i4 = (jj == 3 && o + Ld2 >= Nh) ? o + Ld2 - Nh : o + Ld2;
            sp[jj] = own[o];
          }
          ssc_flip4<ALWAYS1>(own, oth, q, sp, rb, tab);
        }
        __syncthreads();
      }
    }
    return;
  }
  if constexpr (PATH == 6) {
    // N/2 a multiple of 512 and 3 x 128 a multiple of L with an even quotient, Ld2 not
    // dividing 128 (L = 96, 192): site t + 128 k of stream t lies in column
    // (t + 128 k) mod Ld2, which repeats every third k while the row advances by an even
    // number, so the horizontal neighbour's offset from the site depends only on
    // (k mod 3, half).  Vertical neighbours are the site -/+ Ld2, except above the stream's
    // first site (t < Ld2) and below its last (t >= 128 - Ld2).  The offsets of a block's
    // four sites are its base plus constants; six registers hold the horizontal offsets.
    const int Ld2 = d.Ld2, nb = Nh / 512;
    int ec[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const Site q = site_at(t + kStreamsPerUnit * c, h, d);
        ec[h][c] = q.i1 - q.idx;
      }
    // k = 4 (part + P j) + jj, so k mod 3 = (part + P j + jj) mod 3: rotate by part
    const int rot = part % 3;
    int er[2][3];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int y = 0; y < 3; ++y)
        er[h][y] = rot == 0 ? ec[h][y] : rot == 1 ? ec[h][(y + 1) % 3] : ec[h][(y + 2) % 3];
    const int top = t < Ld2 ? Nh : 0;
    const int bot = t >= kStreamsPerUnit - Ld2 ? Nh : 0;
    for (uint32_t s = 0; s < nsw; ++s) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        int8_t* own = lat8 + h * Nh;
        const int8_t* oth = lat8 + (h ^ 1) * Nh;
        for (int j0 = 0; part + P * j0 < nb; j0 += 3) {
#pragma unroll
          for (int jr = 0; jr < 3; ++jr) {
            const int m = part + P * (j0 + jr);
            if (m < nb) {
              const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m);
              const int o0 = t + 512 * m;
              const int up = m == 0 ? top : 0, dn = m == nb - 1 ? bot : 0;
              Site q[4];
              int sp[4];
#pragma unroll
              for (int jj = 0; jj < 4; ++jj) {
                const int o = o0 + jj * kStreamsPerUnit;
                q[jj].idx = o;
                q[jj].i1 = o + er[h][(P * jr + jj) % 3];
                q[jj].i3 = o - Ld2 + (jj == 0 ? up : 0);
                q[jj].i4 = o + Ld2 - (jj == 3 ? dn : 0);
                sp[jj] = own[o];
              }
              ssc_flip4<ALWAYS1>(own, oth, q, sp, rb, tab);
            }
          }
        }
        __syncthreads();
      }
    }
    return;
  }
  if constexpr (PATH == 5) {
    // N/2 not a multiple of 512 and more than 9 sites per stream and half sweep: stream t's
    // n_t sites take the words (2s + h) n_t + k of its stream; the Philox blocks spanning
    // them are dealt to the P threads of the stream, a whole block flips four sites
    // together, the first and last block of a half sweep only the sites it holds words for.
    const int nts = t < Nh ? (Nh - t + kStreamsPerUnit - 1) / kStreamsPerUnit : 0;
    for (uint32_t s = 0; s < nsw; ++s) {
      for (int h = 0; h < 2; ++h) {
        int8_t* own = lat8 + h * Nh;
        const int8_t* oth = lat8 + (h ^ 1) * Nh;
        const uint64_t base = (uint64_t)(2 * s + h) * nts;
        const uint64_t b0 = base >> 2;
        const int nblk = nts ? (int)(((base + nts - 1) >> 2) - b0 + 1) : 0;
        for (int c = part; c < nblk; c += P) {
          const U4 rb = stream_block(seed, sub0 + t, b0 + c);
          const int k0 = (int)(((b0 + c) << 2) - base);
          if (k0 >= 0 && k0 + 3 < nts) {
            Site q[4];
            int sp[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              q[j] = site_at(t + kStreamsPerUnit * (k0 + j), h, d);
              sp[j] = own[q[j].idx];
            }
            ssc_flip4<ALWAYS1>(own, oth, q, sp, rb, tab);
          } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int k = k0 + j;
              if (k >= 0 && k < nts)
                ssc_flip<ALWAYS1>(own, oth, site_at(t + kStreamsPerUnit * k, h, d), u4_word(rb, j), tab);
            }
          }
        }
        __syncthreads();
      }
    }
    return;
  }
  if ((Nh & 511) == 0) {
    // aligned: thread = (stream t, part); blocks m = part, part + P, ... of nb per half
    const int nb = Nh / 512;
    if (PATH == 0 && nb <= P) {
      // at most one block per half sweep per thread: its four sites stay in registers
      const bool act = part < nb;
      Site qs[2][4];
      int sp[2][4];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          qs[h][jj] = site_at(t + kStreamsPerUnit * (4 * part + jj), h, d);
          sp[h][jj] = act ? lat8[h * Nh + qs[h][jj].idx] : 0;
        }
      for (uint32_t s = 0; s < nsw; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (act) {
            int8_t* own = lat8 + h * Nh;
            const int8_t* oth = lat8 + (h ^ 1) * Nh;
            const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + part);
            ssc_flip4<ALWAYS1>(own, oth, qs[h], sp[h], rb, tab);
          }
          __syncthreads();
        }
      }
    } else if (PATH == 1) {
      // up to three blocks per thread and half sweep (L = 96 at 512 threads): the twelve
      // sites' offsets stay in registers, packed two 16-bit offsets per register
      constexpr int MB = 3;
      uint32_t pa[2][MB][4], pb[2][MB][4];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int j = 0; j < MB; ++j)
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            const int m = part + P * j;
            const Site q = site_at(m < nb ? t + kStreamsPerUnit * (4 * m + jj) : 0, h, d);
            pa[h][j][jj] = (uint32_t)q.idx | ((uint32_t)q.i1 << 16);
            pb[h][j][jj] = (uint32_t)q.i3 | ((uint32_t)q.i4 << 16);
          }
      for (uint32_t s = 0; s < nsw; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          int8_t* own = lat8 + h * Nh;
          const int8_t* oth = lat8 + (h ^ 1) * Nh;
#pragma unroll
          for (int j = 0; j < MB; ++j) {
            const int m = part + P * j;
            if (m < nb) {
              const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m);
              Site q[4];
              int sp[4];
#pragma unroll
              for (int jj = 0; jj < 4; ++jj) {
                q[jj].idx = (int)(pa[h][j][jj] & 0xFFFFu);
                q[jj].i1 = (int)(pa[h][j][jj] >> 16);
                q[jj].i3 = (int)(pb[h][j][jj] & 0xFFFFu);
                q[jj].i4 = (int)(pb[h][j][jj] >> 16);
                sp[jj] = own[q[jj].idx];
              }
              ssc_flip4<ALWAYS1>(own, oth, q, sp, rb, tab);
            }
          }
          __syncthreads();
        }
      }
    } else {
      for (uint32_t s = 0; s < nsw; ++s) {
        for (int h = 0; h < 2; ++h) {
          int8_t* own = lat8 + h * Nh;
          const int8_t* oth = lat8 + (h ^ 1) * Nh;
          for (int m = part; m < nb; m += P) {
            const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m);
            const int base = t + kStreamsPerUnit * 4 * m;
            const Site q[4] = {site_at(base, h, d), site_at(base + kStreamsPerUnit, h, d),
                               site_at(base + 2 * kStreamsPerUnit, h, d),
                               site_at(base + 3 * kStreamsPerUnit, h, d)};
            int sp[4] = {own[q[0].idx], own[q[1].idx], own[q[2].idx], own[q[3].idx]};
            ssc_flip4<ALWAYS1>(own, oth, q, sp, rb, tab);
          }
          __syncthreads();
        }
      }
    }
  } else {
    // generic: one thread per stream, sites in stream order with a cached Philox block
    const bool has = threadIdx.x < kStreamsPerUnit && t < Nh;
    const int nt = has ? (Nh - t + kStreamsPerUnit - 1) / kStreamsPerUnit : 0;
    U4 rb{0, 0, 0, 0};
    uint64_t cached = ~0ull;
    for (uint32_t s = 0; s < nsw; ++s) {
      for (int h = 0; h < 2; ++h) {
        int8_t* own = lat8 + h * Nh;
        const int8_t* oth = lat8 + (h ^ 1) * Nh;
        for (int k = 0; k < nt; ++k) {
          const uint64_t p = (uint64_t)(2 * s + h) * nt + k;
          if ((p >> 2) != cached) { cached = p >> 2; rb = stream_block(seed, sub0 + t, cached); }
          ssc_flip<ALWAYS1>(own, oth, site_at(t + kStreamsPerUnit * k, h, d),
                            u4_word(rb, (int)(p & 3)), tab);
        }
        __syncthreads();
      }
    }
  }
}

// SSC: one block per replica.
template <int NT, int PATH>
__global__ __launch_bounds__(NT) void k_sweep_ssc(const RunDesc* __restrict__ runs,
                                                  const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKUnits * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  const int B = blockIdx.x - off[r];
  const int Rn = st->Rnew;
  if (B >= Rn) return;
  extern __shared__ int8_t lat8[];
  const int L = d.L, Ld2 = d.Ld2, Nh = d.Nh, N = d.N;
  const int step = st->step;
  int8_t* g = (int8_t*)d.pop[st->buf ^ 1] + (size_t)B * N;
  for (int i = threadIdx.x; i < N; i += NT) lat8[i] = g[i];
  const uint64_t V1 = d.thr[3 * step + 0], V4 = d.thr[3 * step + 1], V8 = d.thr[3 * step + 2];
  const uint64_t sub0 = st->IS + (uint64_t)st->R + (uint64_t)B * kStreamsPerUnit;
  __shared__ uint32_t tab[9];
  if (threadIdx.x < 9) {
    const int prod = (int)threadIdx.x - 4;
    tab[threadIdx.x] = prod == 4 ? (uint32_t)V8
                     : prod == 2 ? (uint32_t)V4
                                 : (uint32_t)min(V1, (uint64_t)0xFFFFFFFFu);
  }
  __syncthreads();
  if (V1 > 0xFFFFFFFFull)
    ssc_sweeps<NT, true, PATH>(lat8, d, d.seed, sub0, d.eqsweeps, tab);
  else
    ssc_sweeps<NT, false, PATH>(lat8, d, d.seed, sub0, d.eqsweeps, tab);
  // energy and magnetisation (energyKer SSC:195-215)
  int e = 0, mm = 0;
  for (int idx = threadIdx.x; idx < Nh; idx += NT) {
    const int ty = idx / Ld2, tx = idx - ty * Ld2;
    const int iL = ty * Ld2 + (tx == 0 ? Ld2 - 1 : tx - 1);
    const int iU = (ty == L - 1 ? 0 : ty + 1) * Ld2 + tx;
    const int sA = lat8[idx], sB = lat8[Nh + idx];
    int Ai2, Bi2;
    if (ty & 1) { Ai2 = sB; Bi2 = lat8[iL]; } else { Ai2 = lat8[Nh + iL]; Bi2 = sA; }
    const int Ai4 = lat8[Nh + iU], Bi4 = lat8[iU];
    e += -sA * (Ai2 + Ai4) - sB * (Bi2 + Bi4);
    mm += sA + sB;
  }
  __shared__ int shi[16];
  const int E = block_sum_int<int>(e, shi);
  const int Mg = block_sum_int<int>(mm, shi);
  if (threadIdx.x == 0) { d.IE[B] = E; d.M[B] = Mg; }
  for (int i = threadIdx.x; i < N; i += NT) g[i] = lat8[i];
}

// ---- SSC sweeps on packed site bytes (aligned lattices, L <= 256) ----------------------
// The streams of a replica are dealt to threads four at a time: a task is (quad q: streams
// 4q..4q+3, Philox block m), and for each word jj of the block the four sites
// 4q + j + 128 (4m + jj), j = 0..3, are four consecutive sites of one row (L/2 and 4q + 128 k
// are multiples of 4), i.e. one aligned dword of the sublattice in LDS.  Spins are held in
// LDS as bytes 0 / 4, so the four neighbours of all four sites are four dwords (same index,
// above, below, and the horizontal neighbours taken from the adjacent group by one byte
// permute) and the four sites' threshold offsets four bytes of one sum (ssc_dword_update).
//
// Each thread takes MB = 2 tasks (tid and tid + NT): two tasks' eight Philox blocks in flight
// ran 10-30 % faster than one task per thread at equal or higher lane occupancy (L = 96..256),
// four tasks per thread spilled and ran at half speed.

// Bit j of the result = [v_j <= T_j], j = 0..3: compares into lane masks, and an add with
// carry-in per mask (b = 2 b + carry), ordered so that every mask is written four VALU
// instructions before its carry-in read.
__device__ __forceinline__ uint32_t ssc_accept4(uint32_t T0, uint32_t v0, uint32_t T1, uint32_t v1,
                                                uint32_t T2, uint32_t v2, uint32_t T3, uint32_t v3) {
  uint64_t m0, m1, m2, m3, cc;
  uint32_t b;
  asm("v_cmp_ge_u32_e64 %[m3], %[t3], %[v3]\n\t"
      "v_cmp_ge_u32_e64 %[m2], %[t2], %[v2]\n\t"
      "v_cmp_ge_u32_e64 %[m1], %[t1], %[v1]\n\t"
      "v_cmp_ge_u32_e64 %[m0], %[t0], %[v0]\n\t"
      "v_addc_co_u32_e64 %[b], %[cc], 0, 0, %[m3]\n\t"
      "v_addc_co_u32_e64 %[b], %[cc], %[b], %[b], %[m2]\n\t"
      "v_addc_co_u32_e64 %[b], %[cc], %[b], %[b], %[m1]\n\t"
      "v_addc_co_u32_e64 %[b], %[cc], %[b], %[b], %[m0]"
      : [b] "=&v"(b), [m0] "=&s"(m0), [m1] "=&s"(m1), [m2] "=&s"(m2), [m3] "=&s"(m3),
        [cc] "=&s"(cc)
      : [t0] "v"(T0), [v0] "v"(v0), [t1] "v"(T1), [v1] "v"(v1), [t2] "v"(T2), [v2] "v"(v2),
        [t3] "v"(T3), [v3] "v"(v3));
  return b;
}

// Byte selectors of v_perm_b32(Dx, D0, sel) for the horizontal neighbours of a group
// [x, x+3]: the group after (bytes D0[1..3], Dx[0]) or the group before (Dx[3], D0[0..2]).
constexpr uint32_t kSelNext = 0x04030201u, kSelPrev = 0x02010007u;

// LDS byte of a site in the packed-byte SSC kernels: 0 (down) or 4 (up).  Conversion
// from and to the population's int8 +1 / -1 (bit 1 of ~x is the sign of x):
__device__ __forceinline__ uint32_t ssc_to_bytes4(uint32_t g) { return (~g << 1) & 0x04040404u; }
__device__ __forceinline__ uint32_t ssc_from_bytes4(uint32_t x) { return ~((x >> 2) * 0xFEu); }

// Update of one dword = four sites of a row, own bytes S, neighbour bytes D0 (same index),
// Dh (horizontal, after the permute), Du, Dd (rows above and below), Philox words v0..v3.
// With up = 4 the byte sum D0 + Dh + Du + Dd + 5 S = 4 (c + 5 s), c the up neighbours and s
// the own spin, is already the byte offset of the site's threshold class in the table tb (10
// classes, thresholds by the reference's product (2s-1)(2c-4)); no byte exceeds 36.  Per site a
// byte extract, one LDS threshold read, a compare and an add-with-carry that shifts the
// decision into a 4-bit mask; one multiply spreads the mask to bit 2 of the four bytes
// (b x (2^2 + 2^9 + 2^16 + 2^23): bit j lands at 8 j + 2, no carries) to flip them.
// Thresholds are compared as v <= V - 1 (v < V for V >= 1; V = 2^32, every move accepted,
// included); WIDE (a threshold 0, beta above about 2.8) uses the 64-bit form v < V.
template <bool WIDE>
__device__ __forceinline__ uint32_t ssc_dword_update(uint32_t S, uint32_t D0, uint32_t Dh,
                                                     uint32_t Du, uint32_t Dd, uint32_t v0,
                                                     uint32_t v1, uint32_t v2, uint32_t v3,
                                                     const uint8_t* tb, const uint64_t* tabw) {
  // three instructions (the compiler splits 5 S into a shift and an add of its own)
  uint32_t t4;
  asm("v_lshl_add_u32 %[t], %[s], 2, %[d0]\n\t"
      "v_add3_u32 %[t], %[t], %[s], %[du]\n\t"
      "v_add3_u32 %[t], %[t], %[dd], %[dh]"
      : [t] "=&v"(t4)
      : [s] "v"(S), [d0] "v"(D0), [du] "v"(Du), [dd] "v"(Dd), [dh] "v"(Dh));
  uint32_t bits;
  if constexpr (!WIDE) {
    bits = ssc_accept4(*(const uint32_t*)(tb + (t4 & 0xFFu)), v0,
                       *(const uint32_t*)(tb + ((t4 >> 8) & 0xFFu)), v1,
                       *(const uint32_t*)(tb + ((t4 >> 16) & 0xFFu)), v2,
                       *(const uint32_t*)(tb + (t4 >> 24)), v3);
  } else {
    bits = ((uint64_t)v0 < tabw[(t4 & 0xFFu) >> 2] ? 1u : 0u) |
           ((uint64_t)v1 < tabw[((t4 >> 8) & 0xFFu) >> 2] ? 2u : 0u) |
           ((uint64_t)v2 < tabw[((t4 >> 16) & 0xFFu) >> 2] ? 4u : 0u) |
           ((uint64_t)v3 < tabw[(t4 >> 24) >> 2] ? 8u : 0u);
  }
  return S ^ (__umul24(bits, 0x810204u) & 0x04040404u);
}

template <bool B>
struct BoolC {
  static constexpr bool value = B;
};
template <int V>
struct IntC {
  static constexpr int value = V;
};

template <int NTMAX, int MB>
__global__ __launch_bounds__(NTMAX) void k_sweep_ssc_q(const RunDesc* __restrict__ runs,
                                                      const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKUnits * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  const int B = blockIdx.x - off[r];
  if (B >= st->Rnew) return;
  extern __shared__ __align__(16) uint32_t latq[];  // N bytes: sublattice A then B
  const int NT = blockDim.x;
  const int L = d.L, Ld2 = d.Ld2, Nh = d.Nh, N = d.N;
  const int step = st->step;
  uint32_t* g = (uint32_t*)((int8_t*)d.pop[st->buf ^ 1] + (size_t)B * N);
  for (int i = threadIdx.x; i < N / 4; i += NT) latq[i] = ssc_to_bytes4(g[i]);
  const uint64_t V1 = d.thr[3 * step + 0], V4 = d.thr[3 * step + 1], V8 = d.thr[3 * step + 2];
  __shared__ uint32_t tabn[10];
  __shared__ uint64_t tabw[10];
  if (threadIdx.x < 10) {
    const int i = (int)threadIdx.x, up = i >= 5, c = up ? i - 5 : i;
    const int prod = (2 * up - 1) * (2 * c - 4);
    const uint64_t V = prod == 4 ? V8 : prod == 2 ? V4 : V1;
    tabw[i] = V;
    tabn[i] = (uint32_t)(V - 1);
  }
  const bool wide = __builtin_amdgcn_readfirstlane((int)(V4 == 0 || V8 == 0)) != 0;
  const uint64_t seed = d.seed;
  const uint64_t sub0 = st->IS + (uint64_t)st->R + (uint64_t)B * kStreamsPerUnit;
  // the thread's tasks and their fixed geometry (byte offsets within a sublattice)
  const int nb = Nh >> 9, T = 32 * nb;
  if (T > MB * NT) {  // a launch shape that leaves tasks unswept: refuse it (host error bit 2)
    if (threadIdx.x == 0) atomicOr(&st->err, 2);
    return;
  }
  int q[MB], m[MB], o0[MB], up0[MB], dn3[MB];
  int hx[MB][2][4];
  uint32_t sel[MB][2][4];
#pragma unroll
  for (int i = 0; i < MB; ++i) {
    const int task = (int)threadIdx.x + i * NT;
    q[i] = task & 31;
    m[i] = task >> 5;
    o0[i] = 4 * q[i] + 512 * m[i];
    up0[i] = o0[i] < Ld2 ? o0[i] - Ld2 + Nh : o0[i] - Ld2;
    dn3[i] = o0[i] + 384 + Ld2 >= Nh ? o0[i] + 384 + Ld2 - Nh : o0[i] + 384 + Ld2;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int o = o0[i] + 128 * jj;
      const int ty = o / Ld2, tx = o - ty * Ld2;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // A (h = 0): odd rows neighbour tx + 1, even rows tx - 1; B the other way round
        const bool next = ((ty & 1) != 0) == (h == 0);
        hx[i][h][jj] = next ? (tx + 4 < Ld2 ? o + 4 : ty * Ld2) : (tx > 0 ? o - 4 : ty * Ld2 + Ld2 - 4);
        sel[i][h][jj] = next ? kSelNext : kSelPrev;
      }
    }
  }
  __syncthreads();
  const uint8_t* tb = (const uint8_t*)tabn;
  // the threshold form is uniform: one loop per form, no branch per dword
  auto sweeps = [&](auto wide_c) __attribute__((always_inline)) {
    constexpr bool WIDE = decltype(wide_c)::value;
    for (uint32_t s = 0; s < d.eqsweeps; ++s) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        uint8_t* own8 = (uint8_t*)latq + h * Nh;
        const uint8_t* oth8 = (const uint8_t*)latq + (h ^ 1) * Nh;
#pragma unroll
        for (int i = 0; i < MB; ++i) {
          if ((int)threadIdx.x + i * NT >= T) continue;
          const uint64_t blk = (uint64_t)(2 * s + h) * nb + m[i];
          U4 rq[4];
          stream_block_quad(seed, sub0 + 4 * q[i], blk, rq);
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            const int o = o0[i] + 128 * jj;
            const uint32_t S = *(const uint32_t*)(own8 + o);
            const uint32_t D0 = *(const uint32_t*)(oth8 + o);
            const uint32_t Dx = *(const uint32_t*)(oth8 + hx[i][h][jj]);
            const uint32_t Du = *(const uint32_t*)(oth8 + (jj == 0 ? up0[i] : o - Ld2));
            const uint32_t Dd = *(const uint32_t*)(oth8 + (jj == 3 ? dn3[i] : o + Ld2));
            const uint32_t Dh = __builtin_amdgcn_perm(Dx, D0, sel[i][h][jj]);
            *(uint32_t*)(own8 + o) = ssc_dword_update<WIDE>(
                S, D0, Dh, Du, Dd, u4_word(rq[0], jj), u4_word(rq[1], jj), u4_word(rq[2], jj),
                u4_word(rq[3], jj), tb, tabw);
          }
        }
        __syncthreads();
      }
    }
  };
  if (wide) sweeps(BoolC<true>{});
  else sweeps(BoolC<false>{});
  // back to +1 / -1 bytes for the energies (energyKer SSC:195-215) and the store
  for (int i = threadIdx.x; i < N / 4; i += NT) latq[i] = ssc_from_bytes4(latq[i]);
  __syncthreads();
  const int8_t* lat8 = (const int8_t*)latq;
  int e = 0, mm = 0;
  for (int idx = threadIdx.x; idx < Nh; idx += NT) {
    const int ty = idx / Ld2, tx = idx - ty * Ld2;
    const int iL = ty * Ld2 + (tx == 0 ? Ld2 - 1 : tx - 1);
    const int iU = (ty == L - 1 ? 0 : ty + 1) * Ld2 + tx;
    const int sA = lat8[idx], sB = lat8[Nh + idx];
    int Ai2, Bi2;
    if (ty & 1) { Ai2 = sB; Bi2 = lat8[iL]; } else { Ai2 = lat8[Nh + iL]; Bi2 = sA; }
    const int Ai4 = lat8[Nh + iU], Bi4 = lat8[iU];
    e += -sA * (Ai2 + Ai4) - sB * (Bi2 + Bi4);
    mm += sA + sB;
  }
  __shared__ int shi[16];
  const int E = block_sum_int<int>(e, shi);
  const int Mg = block_sum_int<int>(mm, shi);
  if (threadIdx.x == 0) { d.IE[B] = E; d.M[B] = Mg; }
  for (int i = threadIdx.x; i < N / 4; i += NT) g[i] = latq[i];
}

// Packed-byte SSC sweeps of any lattice with L % 8 == 0 (aligned or not, small or large).
// Stream quad q (streams 4q .. 4q+3 of a replica) owns the dwords at byte offsets
// 4q + 128 k, k < n_q = ceil((N/2 - 4q) / 128), of each sublattice: the k-th sites of its four
// streams side by side in one row (L/2 % 4 == 0 keeps a dword inside a row, and N/2 % 4 == 0
// gives the four streams the same n_q).  In half sweep j = 2s + h the quad's k-th dword takes
// word j n_q + k of each of its four streams, so Philox block b of the four streams serves the
// dwords k = 4b + i - j n_q (i < 4) that fall in [0, n_q).  The blocks of a half sweep are
// dealt to P parts of the quad by b mod P; a block that spans two half sweeps stays with its
// part and is evaluated once (the last block is kept in registers).  A workgroup holds
// RPW = NT / (32 P) replicas of one run, each in its own N-byte LDS slot (several small
// lattices per workgroup, or one large lattice over 32 P lanes).  The four dwords of a block
// are straight-line code up to the store: a word outside [0, n_q) reads dword 0 and stores
// nothing.  Site geometry (horizontal neighbour dword and byte
// selector, rows above and below) comes from a per-workgroup table of (k, q) entries when
// P = 1 (at most 12 dwords per quad), else from the dword's offset.  Dword update,
// thresholds and acceptance as in k_sweep_ssc_q.
constexpr uint32_t kSelSum = kSelNext + kSelPrev;

// Geometry of dword o of a sublattice in half h: byte offsets of the horizontal neighbour
// dword, the dwords above and below (other sublattice), and the byte selector.
__device__ __forceinline__ void ssc_p_geo(int o, int h, int L, int Ld2, int Nh, uint32_t magic,
                                          int& hx, int& up, int& dn, uint32_t& sel) {
  const int ty = (int)mulhi32((uint32_t)o, magic);
  const int tx = o - (int)__umul24((uint32_t)ty, (uint32_t)Ld2);
  // A (h = 0): odd rows neighbour tx + 1, even rows tx - 1; B the other way round
  const bool next = ((ty ^ h) & 1) != 0;
  const int hxn = o + 4 - (tx + 4 == Ld2 ? Ld2 : 0);
  const int hxp = o - 4 + (tx == 0 ? Ld2 : 0);
  hx = next ? hxn : hxp;
  sel = next ? kSelNext : kSelPrev;
  up = o - Ld2 + (ty == 0 ? Nh : 0);
  dn = o + Ld2 - (ty == L - 1 ? Nh : 0);
}

template <bool WIDE, bool TAB>
__device__ __forceinline__ void ssc_p_sweeps(uint8_t* lat8, const uint4* geo,
                                             const RunDesc& d, uint32_t nsw, uint64_t seed,
                                             uint64_t sub, int q, int part, int P, int nq,
                                             bool walk, const uint8_t* tb, const uint64_t* tabw) {
  const int L = d.L, Ld2 = d.Ld2, Nh = d.Nh;
  const uint32_t magic = d.ld2_magic;
  U4 c0{0, 0, 0, 0}, c1{0, 0, 0, 0}, c2{0, 0, 0, 0}, c3{0, 0, 0, 0};
  uint32_t cb = 0xFFFFFFFFu;   // block held in c0..c3
  uint32_t b = (uint32_t)part;  // smallest block of this part not yet passed
  for (uint32_t s = 0; s < nsw; ++s) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      uint8_t* own8 = lat8 + h * Nh;
      const uint8_t* oth8 = lat8 + (h ^ 1) * Nh;
      if (walk) {
        const uint32_t w0 = (2 * s + h) * (uint32_t)nq;
        const uint32_t bA = w0 >> 2, bB = (w0 + (uint32_t)nq - 1) >> 2;
        // the previous half sweep's last block may continue into this one
        if (b >= (uint32_t)P && b - (uint32_t)P >= bA) b -= (uint32_t)P;
        for (; b <= bB; b += (uint32_t)P) {
          if (b != cb) {
            U4 rq[4];
            stream_block_quad(seed, sub, b, rq);
            c0 = rq[0];
            c1 = rq[1];
            c2 = rq[2];
            c3 = rq[3];
            cb = b;
          }
          const int k0 = (int)(4 * b - w0);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int k = k0 + i;
            const bool ok = k >= 0 && k < nq;
            // one part: a few dwords per half sweep, skip the words outside it
            if (TAB && !ok) continue;
            const int kk = ok ? k : 0;
            int o, hx, up, dn;
            uint32_t sel;
            if constexpr (TAB) {
              const uint4 e = geo[kk * 32 + q];
              o = (int)e.w;
              hx = h ? (int)(e.x >> 16) : (int)(e.x & 0xFFFFu);
              up = (int)(e.y & 0xFFFFu);
              dn = (int)(e.y >> 16);
              sel = h ? kSelSum - e.z : e.z;
            } else {
              o = 4 * q + 128 * kk;
              ssc_p_geo(o, h, L, Ld2, Nh, magic, hx, up, dn, sel);
            }
            const uint32_t S = *(const uint32_t*)(own8 + o);
            const uint32_t D0 = *(const uint32_t*)(oth8 + o);
            const uint32_t Dx = *(const uint32_t*)(oth8 + hx);
            const uint32_t Du = *(const uint32_t*)(oth8 + up);
            const uint32_t Dd = *(const uint32_t*)(oth8 + dn);
            const uint32_t Dh = __builtin_amdgcn_perm(Dx, D0, sel);
            const uint32_t S2 =
                ssc_dword_update<WIDE>(S, D0, Dh, Du, Dd, u4_word(c0, i), u4_word(c1, i),
                                       u4_word(c2, i), u4_word(c3, i), tb, tabw);
            if (ok) *(uint32_t*)(own8 + o) = S2;
          }
        }
      }
      // one part: a replica's 32 lanes lie in one wavefront, whose LDS operations complete
      // in issue order, so the half sweeps need only a fence against compiler reordering
      if constexpr (TAB) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      } else {
        __syncthreads();
      }
    }
  }
}

// One part and at most one dword per quad and half sweep (N/2 <= 128, L <= 16): half sweep
// j takes word j of the four streams, so each Philox block serves four half sweeps (two
// sweeps) and its word index is known at compile time; the dword's geometry stays in
// registers.  Same update, thresholds and ordering (wavefront-scope fences) as the one-part
// walk of ssc_p_sweeps.
template <bool WIDE>
__device__ __forceinline__ void ssc_p_sweeps_nq1(uint8_t* lat8, const RunDesc& d, uint32_t nsw,
                                                 uint64_t seed, uint64_t sub, int q, bool walk,
                                                 const uint8_t* tb, const uint64_t* tabw) {
  const int L = d.L, Ld2 = d.Ld2, Nh = d.Nh;
  const int o = 4 * q;
  int hx[2], up, dn;
  uint32_t sel[2];
  ssc_p_geo(o, 0, L, Ld2, Nh, d.ld2_magic, hx[0], up, dn, sel[0]);
  ssc_p_geo(o, 1, L, Ld2, Nh, d.ld2_magic, hx[1], up, dn, sel[1]);
  for (uint32_t s = 0; s < nsw; s += 2) {
    U4 c0{0, 0, 0, 0}, c1{0, 0, 0, 0}, c2{0, 0, 0, 0}, c3{0, 0, 0, 0};
    if (walk) {
      U4 rq[4];
      stream_block_quad(seed, sub, s >> 1, rq);
      c0 = rq[0];
      c1 = rq[1];
      c2 = rq[2];
      c3 = rq[3];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (s + (i >> 1) >= nsw) break;  // an odd sweep count ends mid-block (uniform)
      const int h = i & 1;
      uint8_t* own8 = lat8 + h * Nh;
      const uint8_t* oth8 = lat8 + (h ^ 1) * Nh;
      if (walk) {
        const uint32_t S = *(const uint32_t*)(own8 + o);
        const uint32_t D0 = *(const uint32_t*)(oth8 + o);
        const uint32_t Dx = *(const uint32_t*)(oth8 + hx[h]);
        const uint32_t Du = *(const uint32_t*)(oth8 + up);
        const uint32_t Dd = *(const uint32_t*)(oth8 + dn);
        const uint32_t Dh = __builtin_amdgcn_perm(Dx, D0, sel[h]);
        *(uint32_t*)(own8 + o) =
            ssc_dword_update<WIDE>(S, D0, Dh, Du, Dd, u4_word(c0, i), u4_word(c1, i),
                                   u4_word(c2, i), u4_word(c3, i), tb, tabw);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
}

// One part, 257 to 384 sites per sublattice (L = 24): quad q has n_q = C = 3 dwords per half
// sweep for q < qb = (N/2 - 256) / 4 and 2 for the others, so two sweeps (half sweeps
// j = 4J .. 4J + 3) take words 4 J n_q .. 4 J n_q + 4 n_q - 1 of its streams, exactly the
// Philox blocks J n_q + p, p < n_q, and dword k of half sweep hs takes word hs n_q + k of the
// two sweeps: known at compile time per class.  The workgroup's RPW replicas give their
// three-dword quads to the first RPW qb lanes and their two-dword quads to the rest, RPW qb a
// multiple of 64, so every wavefront holds one class: three Philox quads and twelve dword
// updates per two sweeps in the first, two and eight in the others (lanes of both classes in
// one wavefront would each take the larger schedule and select words per class).  A replica
// spans wavefronts, so the half sweeps end in workgroup barriers (outside the class
// branches).  Geometry from the per-workgroup table; update as in k_sweep_ssc_q.
template <bool WIDE>
__device__ __forceinline__ void ssc_p_sweeps_sorted(uint8_t* lat8, const uint4* geo,
                                                    const RunDesc& d, uint32_t nsw, uint64_t seed,
                                                    uint64_t sub, int q, bool big, bool walk,
                                                    const uint8_t* tb, const uint64_t* tabw) {
  const int Nh = d.Nh;
  U4 blk[3][4];  // the class's blocks of the current two sweeps (two used by the 2-dword class)
  auto step = [&](auto c_c, uint32_t s0, int hs) __attribute__((always_inline)) {
    constexpr int C = decltype(c_c)::value;
    const int h = hs & 1;
    uint8_t* own8 = lat8 + h * Nh;
    const uint8_t* oth8 = lat8 + (h ^ 1) * Nh;
#pragma unroll
    for (int p = 0; p < C; ++p)
      if ((4 * p) / C == hs) {
        U4 rq[4];
        stream_block_quad(seed, sub, (uint64_t)(s0 >> 1) * C + p, rq);
#pragma unroll
        for (int j = 0; j < 4; ++j) blk[p][j] = rq[j];
      }
#pragma unroll
    for (int k = 0; k < C; ++k) {
      const int wb = hs * C + k;
      const uint4 e = geo[k * 32 + q];
      const int o = (int)e.w;
      const uint32_t S = *(const uint32_t*)(own8 + o);
      const uint32_t D0 = *(const uint32_t*)(oth8 + o);
      const uint32_t Dx = *(const uint32_t*)(oth8 + (h ? (e.x >> 16) : (e.x & 0xFFFFu)));
      const uint32_t Du = *(const uint32_t*)(oth8 + (e.y & 0xFFFFu));
      const uint32_t Dd = *(const uint32_t*)(oth8 + (e.y >> 16));
      const uint32_t Dh = __builtin_amdgcn_perm(Dx, D0, h ? kSelSum - e.z : e.z);
      *(uint32_t*)(own8 + o) = ssc_dword_update<WIDE>(
          S, D0, Dh, Du, Dd, u4_word(blk[wb >> 2][0], wb & 3), u4_word(blk[wb >> 2][1], wb & 3),
          u4_word(blk[wb >> 2][2], wb & 3), u4_word(blk[wb >> 2][3], wb & 3), tb, tabw);
    }
  };
  for (uint32_t s0 = 0; s0 < nsw; s0 += 2) {
#pragma unroll
    for (int hs = 0; hs < 4; ++hs) {
      if (s0 + (hs >> 1) >= nsw) continue;  // an odd sweep count ends mid-step (uniform)
      if (walk) {
        if (big) step(IntC<3>{}, s0, hs);
        else step(IntC<2>{}, s0, hs);
      }
      __syncthreads();
    }
  }
}

// MODE 0: the stream walk of any L % 8 == 0 (k_sweep_ssc_p); MODE 1: two sweeps per step with
// class-sorted lanes for one part and 257..384 sites per sublattice (k_sweep_ssc_p2, its own
// register budget).
template <int MODE>
__device__ __forceinline__ void ssc_p_body(const RunDesc* __restrict__ runs,
                                           const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKSweep * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  const int NT = blockDim.x;
  const int P = d.ssc_p, lanes = 32 * P, RPW = NT / lanes;
  const int Rn = st->Rnew;
  const int B0 = (blockIdx.x - off[r]) * RPW;
  const int Nh0 = d.Nh;
  if (RPW < 1 || RPW > 32 ||  // a launch shape the plan never makes: refuse (host error bit 2)
      (MODE == 1 && !(P == 1 && Nh0 > 256 && Nh0 <= 384 && Nh0 % 4 == 0 &&
                      (RPW * ((Nh0 - 256) >> 2)) % 64 == 0))) {
    if (threadIdx.x == 0) atomicOr(&st->err, 2);
    return;
  }
  if (B0 >= Rn) return;
  // RPW slots of N bytes (A then B), then for P = 1 the geometry table: 32 n_max entries
  extern __shared__ __align__(16) uint32_t latq[];
  const int L = d.L, Ld2 = d.Ld2, Nh = d.Nh, N = d.N;
  const int slot = threadIdx.x / lanes;  // threads beyond RPW * lanes only join the barriers
  const int lr = threadIdx.x - slot * lanes;
  const int B = B0 + slot;
  const bool live = slot < RPW && B < Rn;
  const int q = lr & 31, part = lr >> 5;
  uint8_t* lat8 = (uint8_t*)latq + (size_t)(live ? slot : 0) * N;
  uint32_t* g = (uint32_t*)((int8_t*)d.pop[st->buf ^ 1] + (size_t)(live ? B : 0) * N);
  if (live)
    for (int i = lr; i < N / 4; i += lanes) ((uint32_t*)lat8)[i] = ssc_to_bytes4(g[i]);
  const bool tab = P == 1;
  uint4* geo = (uint4*)((uint8_t*)latq + (size_t)RPW * N);
  if (tab) {
    const int nmax = (Nh + 127) >> 7;
    for (int i = threadIdx.x; i < 32 * nmax; i += NT) {
      const int qq = i & 31, k = i >> 5, o0 = 4 * qq + 128 * k;
      const int o = o0 < Nh ? o0 : 4 * qq;  // past the quad's last dword: its dword 0
      {
        int hx0, hx1, up, dn;
        uint32_t sel0, sel1;
        ssc_p_geo(o, 0, L, Ld2, Nh, d.ld2_magic, hx0, up, dn, sel0);
        ssc_p_geo(o, 1, L, Ld2, Nh, d.ld2_magic, hx1, up, dn, sel1);
        geo[i] = make_uint4((uint32_t)hx0 | ((uint32_t)hx1 << 16),
                            (uint32_t)up | ((uint32_t)dn << 16), sel0, (uint32_t)o);
      }
    }
  }
  const int step = st->step;
  const uint64_t V1 = d.thr[3 * step + 0], V4 = d.thr[3 * step + 1], V8 = d.thr[3 * step + 2];
  __shared__ uint32_t tabn[10];
  __shared__ uint64_t tabw[10];
  __shared__ int accE[32], accM[32];
  if (threadIdx.x < 10) {
    const int i = (int)threadIdx.x, up = i >= 5, c = up ? i - 5 : i;
    const int prod = (2 * up - 1) * (2 * c - 4);
    const uint64_t V = prod == 4 ? V8 : prod == 2 ? V4 : V1;
    tabw[i] = V;
    tabn[i] = (uint32_t)(V - 1);
  }
  if (threadIdx.x < 32) { accE[threadIdx.x] = 0; accM[threadIdx.x] = 0; }
  const bool wide = __builtin_amdgcn_readfirstlane((int)(V4 == 0 || V8 == 0)) != 0;
  const uint64_t sub = st->IS + (uint64_t)st->R + (uint64_t)B * kStreamsPerUnit + 4 * q;
  const int nq = (Nh - 4 * q + 127) >> 7;
  const bool walk = live && nq > 0;
  __syncthreads();
  const uint8_t* tb = (const uint8_t*)tabn;
  if constexpr (MODE == 1) {
    // sorted lanes: thread i < RPW qb takes three-dword quad i % qb of replica i / qb, the
    // others two-dword quad qb + j % (32 - qb) of replica j / (32 - qb), j = i - RPW qb
    const int qb = (Nh - 256) >> 2, nbig = RPW * qb, tid = (int)threadIdx.x;
    const bool big = tid < nbig;
    const int j = tid - nbig;
    const int slot2 = big ? tid / qb : j / (32 - qb);
    const int q2 = big ? tid % qb : qb + j % (32 - qb);
    const int B2 = B0 + slot2;
    const bool walk2 = tid < RPW * 32 && B2 < Rn;
    uint8_t* lat82 = (uint8_t*)latq + (size_t)(walk2 ? slot2 : 0) * N;
    const uint64_t sub2 = st->IS + (uint64_t)st->R + (uint64_t)B2 * kStreamsPerUnit + 4 * q2;
    if (wide) ssc_p_sweeps_sorted<true>(lat82, geo, d, d.eqsweeps, d.seed, sub2, q2, big, walk2, tb, tabw);
    else ssc_p_sweeps_sorted<false>(lat82, geo, d, d.eqsweeps, d.seed, sub2, q2, big, walk2, tb, tabw);
  } else if (tab && Nh <= 128) {
    if (wide) ssc_p_sweeps_nq1<true>(lat8, d, d.eqsweeps, d.seed, sub, q, walk, tb, tabw);
    else ssc_p_sweeps_nq1<false>(lat8, d, d.eqsweeps, d.seed, sub, q, walk, tb, tabw);
  } else if (wide) {
    if (tab) ssc_p_sweeps<true, true>(lat8, geo, d, d.eqsweeps, d.seed, sub, q, part, P, nq, walk, tb, tabw);
    else ssc_p_sweeps<true, false>(lat8, geo, d, d.eqsweeps, d.seed, sub, q, part, P, nq, walk, tb, tabw);
  } else {
    if (tab) ssc_p_sweeps<false, true>(lat8, geo, d, d.eqsweeps, d.seed, sub, q, part, P, nq, walk, tb, tabw);
    else ssc_p_sweeps<false, false>(lat8, geo, d, d.eqsweeps, d.seed, sub, q, part, P, nq, walk, tb, tabw);
  }
  // back to +1 / -1 bytes for the energies (energyKer SSC:195-215) and the store
  if (live)
    for (int i = lr; i < N / 4; i += lanes) ((uint32_t*)lat8)[i] = ssc_from_bytes4(((uint32_t*)lat8)[i]);
  __syncthreads();
  int e = 0, mm = 0;
  if (live) {
    const int8_t* l8 = (const int8_t*)lat8;
    for (int idx = lr; idx < Nh; idx += lanes) {
      const int ty = idx / Ld2, tx = idx - ty * Ld2;
      const int iL = ty * Ld2 + (tx == 0 ? Ld2 - 1 : tx - 1);
      const int iU = (ty == L - 1 ? 0 : ty + 1) * Ld2 + tx;
      const int sA = l8[idx], sB = l8[Nh + idx];
      int Ai2, Bi2;
      if (ty & 1) { Ai2 = sB; Bi2 = l8[iL]; } else { Ai2 = l8[Nh + iL]; Bi2 = sA; }
      const int Ai4 = l8[Nh + iU], Bi4 = l8[iU];
      e += -sA * (Ai2 + Ai4) - sB * (Bi2 + Bi4);
      mm += sA + sB;
    }
  }
  // integer sums: exact in any order (32-lane groups, then LDS atomics per slot)
#pragma unroll
  for (int m = 16; m >= 1; m >>= 1) {
    e += __shfl_xor(e, m, 32);
    mm += __shfl_xor(mm, m, 32);
  }
  if (live && q == 0) {
    atomicAdd(&accE[slot], e);
    atomicAdd(&accM[slot], mm);
  }
  __syncthreads();
  if (live) {
    if (lr == 0) { d.IE[B] = accE[slot]; d.M[B] = accM[slot]; }
    for (int i = lr; i < N / 4; i += lanes) g[i] = ((const uint32_t*)lat8)[i];
  }
}

__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(6)))
void k_sweep_ssc_p(const RunDesc* __restrict__ runs, const int32_t* __restrict__ offs, int nruns) {
  ssc_p_body<0>(runs, offs, nruns);
}

// Up to 512 threads and 128 VGPRs (four waves per SIMD): the two-sweep step holds the Philox
// quads of its class, which spill inside the sweep loop at 80 VGPRs (six waves, the budget of
// k_sweep_ssc_p) and at 96 (five).
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
void k_sweep_ssc_p2(const RunDesc* __restrict__ runs, const int32_t* __restrict__ offs, int nruns) {
  ssc_p_body<1>(runs, offs, nruns);
}

// MSC sweeps of one unit = (group g, slice qs): sweeps [s_begin, s_end) of the step, with the
// lattice slice staged in LDS.  A slice holds SW = 8*sizeof(S) consecutive replica bits of
// every site word; its LCG chain starts at ran_{q*SW} = A_{q SW} ran_0 + C_{q SW}.  Slices
// let lattices too large for LDS (L=256 x 64 bits = 512 KiB) fit, and do not change any
// stream.  init: the freshly initialised population (buffer 0, energies only, no store).
// FAST: the caller guarantees N/2 % 512 == 0 and nb <= NT / 128 (the register-cached path),
// so the other paths are not compiled in and their registers do not count.
// Ld2 / 2^a for the largest power of two 2^a dividing Ld2 (its odd part): the period in k of
// the column of site t + 128 k when Ld2 < 128.
__device__ __forceinline__ int msc_period(int Ld2) { return Ld2 >> __builtin_ctz((unsigned)Ld2); }

template <typename W, typename S, int NT, bool FAST = false, bool PLAIN = true>
__device__ __attribute__((always_inline)) void msc_unit(const RunDesc& d, const RunState* st, int g, int qs, uint32_t s_begin,
                         uint32_t s_end, bool init, bool energy, S* lat, int* accE, int* accM) {
  constexpr int w = (int)sizeof(W) * 8;
  constexpr int SW = (int)sizeof(S) * 8;
  const int L = d.L, Ld2 = d.Ld2, Nh = d.Nh, N = d.N;
  const int step = st->step;
  uint8_t* gbase = (uint8_t*)d.pop[init ? 0 : st->buf ^ 1] + (size_t)g * N * sizeof(W) +
                   qs * sizeof(S);
  __syncthreads();  // LDS reuse across units
  for (int i = threadIdx.x; i < N; i += NT) lat[i] = *(const S*)(gbase + (size_t)i * sizeof(W));
  const uint32_t T4 = init ? 0u : (uint32_t)d.thr[3 * step + 0];
  const uint32_t T8 = init ? 0u : (uint32_t)d.thr[3 * step + 1];
  const uint64_t sub0 =
      st->IS + (uint64_t)ceil_div(st->R, w) * w + (uint64_t)g * kStreamsPerUnit;
  const uint64_t seed = d.seed;
  // slice qs starts its LCG chain SW * qs steps in; a whole-word slice (SW == w, qs == 0)
  // starts at the stream word itself, known at compile time
  uint32_t JA = 1u, JC = 0u;
  if constexpr (SW != w) lcg_jump(qs * SW, JA, JC);
  __syncthreads();
  if (FAST || (Nh & 511) == 0) {
    constexpr int P = NT / kStreamsPerUnit;
    const int t = threadIdx.x % kStreamsPerUnit;
    const int part = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kStreamsPerUnit));
    const int nb = Nh / 512;
    auto block4 = [&](S* own, const S* oth, const Site (&q)[4], const U4& rb) {
      uint32_t ran[4] = {JA * rb.x + JC, JA * rb.y + JC, JA * rb.z + JC, JA * rb.w + JC};
      S c4[4], c8[4];
      lcg_masks<S, 4>(ran, T4, T8, c4, c8);
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const S sp = own[q[jj].idx];
        own[q[jj].idx] = sp ^ msc_acc<S>(sp, oth[q[jj].i1], oth[q[jj].idx], oth[q[jj].i3],
                                         oth[q[jj].i4], c4[jj], c8[jj]);
      }
    };
    if (FAST || (nb <= P && d.r128 == 0)) {
      // At most one block per half sweep per thread.  N/2 % 512 == 0 and nb <= 8 leave only
      // L = 32 and 64, where L/2 divides 128: the block's four sites lie in column t % Ld2 of
      // rows t / Ld2 + (4 part + jj) q128, so their byte offsets are o0 + 128 B jj with the
      // horizontal delta fixed per (half, jj) and vertical wraps only above jj = 0 and below
      // jj = 3 -- a dozen registers instead of sixteen site indices per half.
      const bool act = part < nb;
      constexpr int B = (int)sizeof(S);
      const int tx = t % Ld2, ty0 = t / Ld2;
      const int dR = (tx + 1 == Ld2 ? 1 - Ld2 : 1) * B, dL = (tx == 0 ? Ld2 - 1 : -1) * B;
      const int dV = Ld2 * B, dN = Nh * B;
      const int o0 = (t + 512 * part) * B, o3b = o0 + 3 * 128 * B;
      const int up0 = o0 < dV ? o0 - dV + dN : o0 - dV;
      const int dn3 = o3b + dV >= dN ? o3b + dV - dN : o3b + dV;
      if constexpr (PLAIN && !FAST) {
        if (nb == 1 && P == 4) {
          // One Philox block per stream and half sweep (L = 32) over four parts: part jj takes
          // the block's word jj, four times the working waves of one part per stream (the rate
          // of these small lattices follows their waves).  The four blocks of two sweeps are
          // evaluated once, one per part, two sweeps ahead during the first half sweep, and
          // handed over in LDS after the lattice slice: two buffers of 4 half sweeps x 4 words
          // x 128 streams (the host adds their 16 KiB to this shape's LDS).
          const int jj = part;
          const int o = (t + 128 * jj) * B;
          const int o3 = jj == 0 ? (t * B < dV ? t * B - dV + dN : t * B - dV) : o - dV;
          const int o4 = jj == 3 ? (o + dV >= dN ? o + dV - dN : o + dV) : o + dV;
          int dh[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) dh[h] = ((ty0 + jj * d.q128 + h) & 1) ? dR : dL;
          uint32_t* wb = (uint32_t*)(lat + N);
          // part p: the block of half sweep 2 g0 + p (sweep g0 + p / 2) into buffer bf
          auto fill = [&](uint32_t g0, int bf) __attribute__((always_inline)) {
            const uint32_t sp = g0 + (uint32_t)(part >> 1);
            if (sp >= s_begin && sp < s_end) {
              const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * g0 + part));
              uint32_t* q = wb + (size_t)((bf * 4 + part) * 4) * kStreamsPerUnit + t;
              q[0] = rb.x;
              q[kStreamsPerUnit] = rb.y;
              q[2 * kStreamsPerUnit] = rb.z;
              q[3 * kStreamsPerUnit] = rb.w;
            }
          };
          const uint32_t first = s_begin & ~1u;
          fill(first, 0);
          __syncthreads();
          int bf = 0;
          for (uint32_t s0 = first; s0 < s_end; s0 += 2, bf ^= 1) {
#pragma unroll
            for (int hs = 0; hs < 4; ++hs) {
              // the next two sweeps' blocks go to the other buffer, last read two sweeps ago
              if (hs == 0 && s0 + 2 < s_end) fill(s0 + 2, bf ^ 1);
              const uint32_t s = s0 + (hs >> 1);
              if (s < s_begin || s >= s_end) continue;  // uniform
              const int h = hs & 1;
              uint8_t* own = (uint8_t*)(lat + h * Nh);
              const uint8_t* oth = (const uint8_t*)(lat + (h ^ 1) * Nh);
              const uint32_t word = wb[(size_t)((bf * 4 + hs) * 4 + jj) * kStreamsPerUnit + t];
              uint32_t ran[1] = {JA * word + JC};
              S c4[1], c8[1];
              if constexpr (SW >= 16) lcg_masks_split<S>(ran[0], T4, T8, c4[0], c8[0]);
              else lcg_masks<S, 1>(ran, T4, T8, c4, c8);
              const S sp = *(const S*)(own + o);
              *(S*)(own + o) = sp ^ msc_acc<S>(sp, *(const S*)(oth + o + dh[h]),
                                               *(const S*)(oth + o), *(const S*)(oth + o3),
                                               *(const S*)(oth + o4), c4[0], c8[0]);
              __syncthreads();
            }
          }
          goto unit_done;
        }
      }
      int d1[2][4];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) d1[h][jj] = ((ty0 + jj * d.q128 + h) & 1) ? dR : dL;
      for (uint32_t s = s_begin; s < s_end; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          if (act) {
            uint8_t* own = (uint8_t*)(lat + h * Nh);
            const uint8_t* oth = (const uint8_t*)(lat + (h ^ 1) * Nh);
            const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + part);
            uint32_t ran[4] = {JA * rb.x + JC, JA * rb.y + JC, JA * rb.z + JC, JA * rb.w + JC};
            S c4[4], c8[4];
            lcg_masks<S, 4>(ran, T4, T8, c4, c8);
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
              const int o = o0 + jj * 128 * B;
              const int o3 = jj == 0 ? up0 : o - dV, o4 = jj == 3 ? dn3 : o + dV;
              const S sp = *(const S*)(own + o);
              *(S*)(own + o) = sp ^ msc_acc<S>(sp, *(const S*)(oth + o + d1[h][jj]),
                                               *(const S*)(oth + o), *(const S*)(oth + o3),
                                               *(const S*)(oth + o4), c4[jj], c8[jj]);
            }
          }
          __syncthreads();
        }
      }
    } else if (d.r128 == 0) {
      // Ld2 divides 128: every site of thread t lies in column t % Ld2, and site
      // t + 128 k sits in row t / Ld2 + k q128.  Row parities within a block repeat from
      // block to block, so the horizontal neighbour's offset is fixed per (half, jj); the
      // vertical neighbours wrap only above jj = 0 (row 0) and below jj = 3 (row L - 1).
      // All offsets are bytes into the LDS slice.
      constexpr int B = (int)sizeof(S);
      const int tx = t % Ld2, ty0 = t / Ld2;
      const int dR = (tx + 1 == Ld2 ? 1 - Ld2 : 1) * B, dL = (tx == 0 ? Ld2 - 1 : -1) * B;
      const int dV = Ld2 * B, dN = Nh * B;
      int d1[2][4];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) d1[h][jj] = ((ty0 + jj * d.q128 + h) & 1) ? dR : dL;
      for (uint32_t s = s_begin; s < s_end; ++s) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          uint8_t* own = (uint8_t*)(lat + h * Nh);
          const uint8_t* oth = (const uint8_t*)(lat + (h ^ 1) * Nh);
          for (int m = part; m < nb; m += P) {
            const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m);
            uint32_t ran[4] = {JA * rb.x + JC, JA * rb.y + JC, JA * rb.z + JC, JA * rb.w + JC};
            S c4[4], c8[4];
            lcg_masks<S, 4>(ran, T4, T8, c4, c8);
            const int o0 = (t + 512 * m) * B;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
              const int o = o0 + jj * 128 * B;
              int o3 = o - dV, o4 = o + dV;
              if (jj == 0 && o < dV) o3 += dN;
              if (jj == 3 && o4 >= dN) o4 -= dN;
              const S sp = *(const S*)(own + o);
              *(S*)(own + o) = sp ^ msc_acc<S>(sp, *(const S*)(oth + o + d1[h][jj]),
                                               *(const S*)(oth + o), *(const S*)(oth + o3),
                                               *(const S*)(oth + o4), c4[jj], c8[jj]);
            }
          }
          __syncthreads();
        }
      }
    } else if (Ld2 < 128 && (msc_period(Ld2) == 3 ||
                             (PLAIN && sizeof(W) == 4 && (msc_period(Ld2) == 5 || msc_period(Ld2) == 7)))) {
      // Ld2 = PER 2^a below 128 with PER = 3, 5, 7 (L = 96, 192; 160; 224; 5 and 7 for 32-bit
      // words in the plain sweep kernel only: the unrolled PER-block steps are long, and the
      // balanced kernel, whose registers the headline shape uses to the limit, does not carry
      // them; these lattices' 32-bit slices exceed 80 KiB and take 1024-thread plain
      // workgroups): site t + 128 k of
      // stream t lies in column (t + 128 k) mod Ld2, which repeats every PER-th k while the row
      // advances by 128 PER / Ld2 = 2^(7-a), an even number, so the horizontal neighbour's
      // offset depends only on (k mod PER, half); k = 4 (part + P j) + jj with j a multiple of
      // PER plus jr gives k mod PER = (4 part + 4 P jr + jj) mod PER.  Vertical neighbours wrap
      // only above the stream's first site and below its last, which holds only while
      // Ld2 < 128 (at L = 384 the first two sites of a stream share row 0).
      auto period = [&](auto per_c) __attribute__((always_inline)) {
        constexpr int PER = decltype(per_c)::value;
        constexpr int B = (int)sizeof(S);
        int ec[2][PER];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int c = 0; c < PER; ++c) {
            const Site q = site_at(t + kStreamsPerUnit * c, h, d);
            ec[h][c] = (q.i1 - q.idx) * B;
          }
        const int rot = (4 * part) % PER;
        int er[2][PER];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int y = 0; y < PER; ++y) {
            int v = ec[h][y];
#pragma unroll
            for (int r = 1; r < PER; ++r)
              if (rot == r) v = ec[h][(y + r) % PER];
            er[h][y] = v;
          }
        const int dV = Ld2 * B, dN = Nh * B;
        const int top = t < Ld2 ? dN : 0, bot = t >= kStreamsPerUnit - Ld2 ? dN : 0;
        for (uint32_t s = s_begin; s < s_end; ++s) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            uint8_t* own = (uint8_t*)(lat + h * Nh);
            const uint8_t* oth = (const uint8_t*)(lat + (h ^ 1) * Nh);
            for (int j0 = 0; part + P * j0 < nb; j0 += PER) {
#pragma unroll
              for (int jr = 0; jr < PER; ++jr) {
                const int m = part + P * (j0 + jr);
                if (m < nb) {
                  const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m);
                  uint32_t ran[4] = {JA * rb.x + JC, JA * rb.y + JC, JA * rb.z + JC, JA * rb.w + JC};
                  S c4[4], c8[4];
                  lcg_masks<S, 4>(ran, T4, T8, c4, c8);
                  const int o0 = (t + 512 * m) * B;
                  const int up = m == 0 ? top : 0, dn = m == nb - 1 ? bot : 0;
#pragma unroll
                  for (int jj = 0; jj < 4; ++jj) {
                    const int o = o0 + jj * 128 * B;
                    const int o1 = o + er[h][(4 * P * jr + jj) % PER];
                    const int o3 = o - dV + (jj == 0 ? up : 0);
                    const int o4 = o + dV - (jj == 3 ? dn : 0);
                    const S sp = *(const S*)(own + o);
                    *(S*)(own + o) = sp ^ msc_acc<S>(sp, *(const S*)(oth + o1), *(const S*)(oth + o),
                                                     *(const S*)(oth + o3), *(const S*)(oth + o4),
                                                     c4[jj], c8[jj]);
                  }
                }
              }
            }
            __syncthreads();
          }
        }
      };
      const int per = msc_period(Ld2);
      if constexpr (PLAIN && sizeof(W) == 4) {
        if (per == 3) period(IntC<3>{});
        else if (per == 5) period(IntC<5>{});
        else period(IntC<7>{});
      } else {
        period(IntC<3>{});
      }
    } else {
      for (uint32_t s = s_begin; s < s_end; ++s) {
        for (int h = 0; h < 2; ++h) {
          for (int m = part; m < nb; m += P) {
            const int base = t + kStreamsPerUnit * 4 * m;
            const Site q[4] = {site_at(base, h, d), site_at(base + kStreamsPerUnit, h, d),
                               site_at(base + 2 * kStreamsPerUnit, h, d),
                               site_at(base + 3 * kStreamsPerUnit, h, d)};
            block4(lat + h * Nh, lat + (h ^ 1) * Nh, q,
                   stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m));
          }
          __syncthreads();
        }
      }
    }
  } else if (Nh <= kStreamsPerUnit) {
    // At most one site per stream and half sweep (L <= 16): half sweep j = 2s + h takes word
    // j of stream t, so one Philox block serves two sweeps with compile-time word indices, and
    // the two sites' neighbour indices stay in registers.  Threads t >= N/2 (and those of
    // other parts) only join the barriers.
    const int t = (int)threadIdx.x;
    const bool act = t < Nh;
    const Site qh[2] = {site_at(act ? t : 0, 0, d), site_at(act ? t : 0, 1, d)};
    for (uint32_t s0 = s_begin & ~1u; s0 < s_end; s0 += 2) {
      U4 rb{0, 0, 0, 0};
      if (act) rb = stream_block(seed, sub0 + t, s0 >> 1);
#pragma unroll
      for (int hs = 0; hs < 4; ++hs) {
        const uint32_t s = s0 + (hs >> 1);
        if (s < s_begin || s >= s_end) continue;  // uniform: a range starting or ending mid-block
        const int h = hs & 1;
        if (act) {
          S* own = lat + h * Nh;
          const S* oth = lat + (h ^ 1) * Nh;
          uint32_t ran[1] = {JA * u4_word(rb, hs) + JC};
          S c4[1], c8[1];
          if constexpr (SW >= 16) lcg_masks_split<S>(ran[0], T4, T8, c4[0], c8[0]);
          else lcg_masks<S, 1>(ran, T4, T8, c4, c8);
          const Site& q = qh[h];
          const S sp = own[q.idx];
          own[q.idx] = sp ^ msc_acc<S>(sp, oth[q.i1], oth[q.idx], oth[q.i3], oth[q.i4], c4[0], c8[0]);
        }
        __syncthreads();
      }
    }
  } else if (Nh <= 9 * kStreamsPerUnit) {
    // At most 9 sites per stream and half sweep (L <= 48), words dealt, not blocks: stream t's
    // n_t sites take the consecutive words (2s + h) n_t + k, k < n_t, of its stream, and word
    // k goes to thread k mod P of the stream (P = NT / 128).  Each thread walks its words in
    // order with the last Philox block cached (a block whose words go to several threads is
    // evaluated by each of them), and a word's LCG values run as four jump-started chains, so
    // no word outside the half sweep is evaluated: dealing whole blocks computes and drops
    // 25-33 % of the words here (R = 20000, attempts/ns: L = 40 2443 -> 2760, 48 2523 -> 2951;
    // at L >= 56 the repeated blocks cost more than the dropped words saved).
    constexpr int P = NT / kStreamsPerUnit;
    const int t = threadIdx.x % kStreamsPerUnit;
    const int part = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kStreamsPerUnit));
    const int nts = t < Nh ? (Nh - t + kStreamsPerUnit - 1) / kStreamsPerUnit : 0;
    U4 rb{0, 0, 0, 0};
    uint64_t cached = ~0ull;
    for (uint32_t s = s_begin; s < s_end; ++s) {
      for (int h = 0; h < 2; ++h) {
        S* own = lat + h * Nh;
        const S* oth = lat + (h ^ 1) * Nh;
        const uint64_t base = (uint64_t)(2 * s + h) * nts;
        for (int k = part; k < nts; k += P) {
          const uint64_t p = base + k;
          if ((p >> 2) != cached) { cached = p >> 2; rb = stream_block(seed, sub0 + t, cached); }
          uint32_t ran[1] = {JA * u4_word(rb, (int)(p & 3)) + JC};
          S c4[1], c8[1];
          // one word in flight: its LCG values as four jump-started chains (16-bit slices
          // and wider)
          if constexpr (SW >= 16) lcg_masks_split<S>(ran[0], T4, T8, c4[0], c8[0]);
          else lcg_masks<S, 1>(ran, T4, T8, c4, c8);
          const Site q = site_at(t + kStreamsPerUnit * k, h, d);
          const S sp = own[q.idx];
          own[q.idx] = sp ^ msc_acc<S>(sp, oth[q.i1], oth[q.idx], oth[q.i3], oth[q.i4], c4[0], c8[0]);
        }
        __syncthreads();
      }
    }
  } else {
    // Half sweeps that do not split into whole Philox blocks: stream t's n_t sites take the
    // consecutive words (2s + h) n_t + k, k < n_t, of its stream.  The blocks spanning those
    // words are dealt round robin to the P = NT / 128 threads of the stream; each block's
    // four words drive four independent LCG chains, and a word outside the half sweep
    // (first or last block) is computed and dropped.
    constexpr int P = NT / kStreamsPerUnit;
    const int t = threadIdx.x % kStreamsPerUnit, part = threadIdx.x / kStreamsPerUnit;
    const int nts = t < Nh ? (Nh - t + kStreamsPerUnit - 1) / kStreamsPerUnit : 0;
    for (uint32_t s = s_begin; s < s_end; ++s) {
      for (int h = 0; h < 2; ++h) {
        S* own = lat + h * Nh;
        const S* oth = lat + (h ^ 1) * Nh;
        const uint64_t base = (uint64_t)(2 * s + h) * nts;
        const uint64_t b0 = base >> 2;
        const int nblk = nts ? (int)(((base + nts - 1) >> 2) - b0 + 1) : 0;
        for (int c = part; c < nblk; c += P) {
          const U4 rb = stream_block(seed, sub0 + t, b0 + c);
          uint32_t ran[4] = {JA * rb.x + JC, JA * rb.y + JC, JA * rb.z + JC, JA * rb.w + JC};
          S c4[4], c8[4];
          lcg_masks<S, 4>(ran, T4, T8, c4, c8);
          const int k0 = (int)(((b0 + c) << 2) - base);
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int k = k0 + j;
            if (k >= 0 && k < nts) {
              const Site q = site_at(t + kStreamsPerUnit * k, h, d);
              const S sp = own[q.idx];
              own[q.idx] = sp ^ msc_acc<S>(sp, oth[q.i1], oth[q.idx], oth[q.i3], oth[q.i4],
                                           c4[j], c8[j]);
            }
          }
        }
        __syncthreads();
      }
    }
  }
unit_done:
  if (energy) msc_energy_slice<S>(lat, L, Ld2, Nh, accE, accM, qs * SW, d.IE, d.M, (size_t)g * w);
  if (!init) {
    __syncthreads();
    for (int i = threadIdx.x; i < N; i += NT) *(S*)(gbase + (size_t)i * sizeof(W)) = lat[i];
  }
}

// msc_energy_slice for a unit whose sublattice A is in HBM (the unit's slice of every site
// word, stride sizeof(W)) and whose sublattice B is in LDS.
template <typename W, typename S>
__device__ void msc_energy_half(const uint8_t* gA, const S* latB, int L, int Ld2, int Nh,
                                int* accE, int* accM, int lane0, int32_t* IE, int32_t* Mo,
                                size_t lane_base) {
  constexpr int SW = (int)sizeof(S) * 8;
  auto A = [&](int i) { return *(const S*)(gA + (size_t)i * sizeof(W)); };
  for (int i = threadIdx.x; i < SW; i += blockDim.x) { accE[i] = 0; accM[i] = 0; }
  __syncthreads();
  const int nq = blockDim.x / SW > 0 ? blockDim.x / SW : 1;
  const int lane = threadIdx.x % SW, q = threadIdx.x / SW;
  if (q < nq) {
    int e = 0, m = 0;
    for (int idx = q; idx < Nh; idx += nq) {
      const int ty = idx / Ld2, tx = idx - ty * Ld2;
      const int iL = ty * Ld2 + (tx == 0 ? Ld2 - 1 : tx - 1);
      const int iU = (ty == L - 1 ? 0 : ty + 1) * Ld2 + tx;
      const S sA = A(idx), sB = latB[idx];
      S Ai2, Bi2;
      if (ty & 1) { Ai2 = sB; Bi2 = A(iL); } else { Ai2 = latB[iL]; Bi2 = sA; }
      const S Ai4 = latB[iU], Bi4 = A(iU);
      const S I1 = sA ^ Ai2, I2 = sA ^ Ai4, I3 = sB ^ Bi2, I4 = sB ^ Bi4;
      const int c = (int)((I1 >> lane) & 1) + (int)((I2 >> lane) & 1) + (int)((I3 >> lane) & 1) +
                    (int)((I4 >> lane) & 1);
      e += 2 * c - 4;
      m += 2 * ((int)((sA >> lane) & 1) + (int)((sB >> lane) & 1)) - 2;
    }
    atomicAdd(&accE[lane], e);
    atomicAdd(&accM[lane], m);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SW; i += blockDim.x) {
    IE[lane_base + lane0 + i] = accE[i];
    Mo[lane_base + lane0 + i] = accM[i];
  }
}

// MSC sweeps of one unit whose 32-bit slice is too large for LDS as a whole (L = 256:
// 256 KiB): LDS holds one sublattice (N/2 words) and the other lives in the registers of the
// threads that update it.  Half sweep h updates the register-resident sublattice h reading
// its neighbours from LDS (sublattice h ^ 1); then every thread swaps its sites' words with
// the LDS copy, so the next half sweep finds h in LDS and h ^ 1 in registers.  A site's
// thread is the one that evaluates its Philox block; the sites lie on the thread's column
// (L/2 divides 128, N/2 % 512 == 0, at most 8 blocks per thread: the caller guarantees it).
// Twice the replica bits per LDS word of the 16-bit slices that would fit otherwise, so
// the per-word work is spread over twice the attempts.
template <typename W>
__device__ __attribute__((always_inline)) void msc_unit_half(const RunDesc& d, const RunState* st, int g, int qs, bool init,
                              uint32_t* lat, int* accE, int* accM) {
  using S = uint32_t;
  constexpr int NT = 1024, P = NT / kStreamsPerUnit, MAXB = 8;
  constexpr int w = (int)sizeof(W) * 8;
  constexpr int SW = 32;
  const int L = d.L, Ld2 = d.Ld2, Nh = d.Nh, N = d.N;
  const int step = st->step;
  const uint32_t nsw = init ? 0u : d.eqsweeps;
  uint8_t* gbase = (uint8_t*)d.pop[init ? 0 : st->buf ^ 1] + (size_t)g * N * sizeof(W) +
                   qs * sizeof(S);
  const int t = threadIdx.x % kStreamsPerUnit, part = threadIdx.x / kStreamsPerUnit;
  const int nb = Nh / 512;
  const int nbt = part < nb ? (nb - part + P - 1) / P : 0;
  __syncthreads();  // LDS reuse across units
  S own[MAXB][4];
#pragma unroll
  for (int i = 0; i < MAXB; ++i)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const int o = t + 512 * (part + P * i) + kStreamsPerUnit * jj;
      own[i][jj] = i < nbt ? *(const S*)(gbase + (size_t)o * sizeof(W)) : 0u;
    }
  for (int i = threadIdx.x; i < Nh; i += NT) lat[i] = *(const S*)(gbase + (size_t)(Nh + i) * sizeof(W));
  const uint32_t T4 = init ? 0u : (uint32_t)d.thr[3 * step + 0];
  const uint32_t T8 = init ? 0u : (uint32_t)d.thr[3 * step + 1];
  const uint64_t sub0 =
      st->IS + (uint64_t)ceil_div(st->R, w) * w + (uint64_t)g * kStreamsPerUnit;
  const uint64_t seed = d.seed;
  uint32_t JA = 1u, JC = 0u;
  if constexpr (SW != w) lcg_jump(qs * SW, JA, JC);
  const int tx = t % Ld2, ty0 = t / Ld2;
  const int dR = tx + 1 == Ld2 ? 1 - Ld2 : 1, dL = tx == 0 ? Ld2 - 1 : -1;
  int d1[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) d1[h][jj] = ((ty0 + jj * d.q128 + h) & 1) ? dR : dL;
  __syncthreads();
  for (uint32_t s = 0; s < nsw; ++s) {
    for (int h = 0; h < 2; ++h) {
      // the sites' offsets are recomputed in every half sweep from an opaque base: held
      // across the loop, the 32 sites' neighbour offsets would not fit the registers
      int obase = t + 512 * part;
      asm volatile("" : "+v"(obase));
#pragma unroll
      for (int i = 0; i < MAXB; ++i) {
        if (i < nbt) {
          const int m = part + P * i;
          const U4 rb = stream_block(seed, sub0 + t, (uint64_t)(2 * s + h) * nb + m);
          uint32_t ran[4] = {JA * rb.x + JC, JA * rb.y + JC, JA * rb.z + JC, JA * rb.w + JC};
          S c4[4], c8[4];
          lcg_masks<S, 4>(ran, T4, T8, c4, c8);
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            const int o = obase + 512 * P * i + kStreamsPerUnit * jj;
            const int o3 = (jj == 0 && o < Ld2) ? o - Ld2 + Nh : o - Ld2;
            const int o4 = (jj == 3 && o + Ld2 >= Nh) ? o + Ld2 - Nh : o + Ld2;
            const int dh = h ? d1[1][jj] : d1[0][jj];
            own[i][jj] ^= msc_acc<S>(own[i][jj], lat[o + dh], lat[o], lat[o3], lat[o4], c4[jj], c8[jj]);
          }
        }
      }
      __syncthreads();
      asm volatile("" : "+v"(obase));
#pragma unroll
      for (int i = 0; i < MAXB; ++i)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
          if (i < nbt) {
            const int o = obase + 512 * P * i + kStreamsPerUnit * jj;
            const S x = lat[o];
            lat[o] = own[i][jj];
            own[i][jj] = x;
          }
      __syncthreads();
    }
  }
  // an even number of swaps: registers hold sublattice A, LDS sublattice B
  if (!init) {
#pragma unroll
    for (int i = 0; i < MAXB; ++i)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
        if (i < nbt) {
          const int o = t + 512 * (part + P * i) + kStreamsPerUnit * jj;
          *(S*)(gbase + (size_t)o * sizeof(W)) = own[i][jj];
        }
    for (int i = threadIdx.x; i < Nh; i += NT) *(S*)(gbase + (size_t)(Nh + i) * sizeof(W)) = lat[i];
  }
  // energies read A back from HBM: the agent-scope fence completes this workgroup's stores and
  // drops stale L1 lines before any wave reads them
  __threadfence();
  __syncthreads();
  msc_energy_half<W, S>(gbase, lat, L, Ld2, Nh, accE, accM, qs * SW, d.IE, d.M, (size_t)g * w);
}

template <typename W, bool INIT>
__global__ __launch_bounds__(1024) void k_sweep_msc_half(const RunDesc* __restrict__ runs,
                                                         const int32_t* __restrict__ offs, int nruns) {
  constexpr int w = (int)sizeof(W) * 8;
  constexpr int NS = w / 32;
  const int* off = offs + kKUnits * (nruns + 1);
  const int bu = blockIdx.x / NS;
  const int qs = blockIdx.x % NS;
  const int r = find_run(off, nruns, bu);
  const RunDesc& d = runs[r];
  const RunState* st = d.st;
  if (!INIT && (!st->active || st->overflow)) return;
  const int g = bu - off[r];
  const int Rn = INIT ? st->R : st->Rnew;
  if (g >= ceil_div(Rn, w)) return;
  extern __shared__ __align__(16) uint8_t smem_sw[];
  __shared__ int accE[32], accM[32];
  msc_unit_half<W>(d, st, g, qs, INIT, (uint32_t*)smem_sw, accE, accM);
}

// MSC: one block per (group, slice), all EqSw sweeps of the step.
template <typename W, typename S, int NT, bool INIT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_sweep_msc(const RunDesc* __restrict__ runs,
                                                  const int32_t* __restrict__ offs, int nruns) {
  constexpr int w = (int)sizeof(W) * 8;
  constexpr int SW = (int)sizeof(S) * 8;
  constexpr int NS = w / SW;  // slices per word
  const int* off = offs + kKUnits * (nruns + 1);
  const int bu = blockIdx.x / NS;
  const int qs = blockIdx.x % NS;
  const int r = find_run(off, nruns, bu);
  const RunDesc& d = runs[r];
  const RunState* st = d.st;
  if (!INIT && (!st->active || st->overflow)) return;
  const int g = bu - off[r];
  const int Rn = INIT ? st->R : st->Rnew;
  if (g >= ceil_div(Rn, w)) return;
  extern __shared__ __align__(16) uint8_t smem_sw[];
  __shared__ int accE[SW], accM[SW];
  msc_unit<W, S, NT>(d, st, g, qs, 0, INIT ? 0u : d.eqsweeps, INIT, true, (S*)smem_sw, accE, accM);
}

// Makespan-balanced MSC sweeps: P co-resident workers split the step's work line
// (units x sweeps, weighted by lattice size) into equal shares, McNaughton's wrap-around
// rule.  A unit cut by a share boundary at sweep k runs sweeps [0, k) first, at the start of
// the later worker, and sweeps [k, E) last, at the end of the earlier worker; the lattice is
// handed over through HBM with an agent-scope release/acquire flag tagged with the step.
// Since a unit's work never exceeds a share, its two parts never overlap in time and the
// wait is normally empty; it is bounded, and a timeout is reported as an error.
__device__ __forceinline__ int64_t run_units(const RunDesc& d, int NS) {
  const RunState* st = d.st;
  if (!st->active || st->overflow) return 0;
  return (int64_t)ceil_div(st->Rnew, d.w) * NS;
}

// At most 128 VGPRs (4 waves per SIMD), so two 512-thread workers fit a CU; the few
// registers this spills are cheaper than running one worker per CU (7 % at L=64).
template <typename W, typename S, int NT, bool FAST>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(FAST ? 6 : 4))) void k_sweep_msc_bal(const RunDesc* __restrict__ runs,
                                                      int nruns, int P) {
  constexpr int SW = (int)sizeof(S) * 8;
  constexpr int NS = (int)sizeof(W) * 8 / SW;
  extern __shared__ __align__(16) uint8_t smem_sw[];
  __shared__ int accE[SW], accM[SW];
  S* lat = (S*)smem_sw;
  const int wk = blockIdx.x;
  int64_t Ctot = 0, Umax = 0;
  for (int r = 0; r < nruns; ++r) {
    const int64_t uc = (int64_t)runs[r].eqsweeps * runs[r].N;
    const int64_t n = run_units(runs[r], NS);
    Ctot += n * uc;
    if (n > 0 && uc > Umax) Umax = uc;
  }
  if (Ctot == 0) return;
  // a share never smaller than the largest unit, so a unit is cut at most once
  const int64_t Sh = max((Ctot + P - 1) / P, Umax);
  const int64_t a = (int64_t)wk * Sh, b = min(a + Sh, Ctot);
  if (a >= b) return;
  __shared__ int timed_out;
  int64_t base = 0;
  for (int r = 0; r < nruns && base < b; ++r) {
    const RunDesc& d = runs[r];
    const int64_t E = d.eqsweeps, uc = E * d.N;
    const int64_t n = run_units(d, NS);
    const int64_t cr = n * uc;
    if (base + cr <= a) { base += cr; continue; }
    const int64_t u0 = a > base ? (a - base) / uc : 0;
    const int64_t u1 = min(n - 1, (b - 1 - base) / uc);
    RunState* st = d.st;
    const uint32_t tag = (uint32_t)st->step;
    for (int64_t u = u0; u <= u1; ++u) {
      const int64_t p0 = base + u * uc;
      const int g = (int)(u / NS), qs = (int)(u % NS);
      uint32_t* flag = d.flags + (size_t)g * NS + qs;
      // task: sweeps [s0, s1); publish = hand the lattice to the earlier share afterwards;
      // wait = take it over from the later share first
      int64_t s0 = 0, s1 = E;
      bool publish = false, wait = false;
      if (p0 < a) {  // first sweeps of a unit that began in the previous share
        const int64_t kprev = (a - p0 + d.N / 2) / d.N;
        s1 = E - kprev;
        publish = kprev > 0;
      } else if (p0 + uc > b) {  // last sweeps of a unit that continues in the next share
        const int64_t kme = (b - p0 + d.N / 2) / d.N;
        s0 = E - kme;
        wait = kme < E;
      }
      if (s1 <= s0) continue;
      if (wait) {
        if (threadIdx.x == 0) {
          timed_out = 0;
          uint32_t spins = 0;
          while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != tag) {
            __builtin_amdgcn_s_sleep(8);
            if (++spins > (1u << 26)) { timed_out = 1; break; }
          }
          if (timed_out) atomicOr(&st->err, 1);
        }
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      }
      msc_unit<W, S, NT, FAST, false>(d, st, g, qs, (uint32_t)s0, (uint32_t)s1, false, !publish, lat, accE, accM);
      if (publish) {
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(flag, tag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    base += cr;
  }
}

// ============================================================================================
// Averages (CalcAverages SSC:282-292 / MSC:437-450): block tree sums of E, E^2, |M|, M^2,
// M^4 over 1024 replicas (SSC) or over 1024 groups for one lane (MSC: block = (chunk, lane)),
// plus integer sums.  The run's last block adds the block partials in launch order and
// writes the step record (finish_step).
// ============================================================================================
__device__ void finish_step(const RunDesc& d, RunState* st, int nblk);

__global__ __launch_bounds__(1024) void k_avg(const RunDesc* __restrict__ runs,
                                              const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKAvg * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!st->active || st->overflow) return;
  const int c = blockIdx.x - off[r];
  const int Rn = st->Rnew, w = d.w, buf = st->buf ^ 1;
  int nb, b, by, nblk;
  if (d.coding == kSSC) {
    nb = ceil_div(Rn, kRB);
    nblk = nb;
    if (c >= nb) return;
    b = c; by = 0;
  } else {
    nb = ceil_div(ceil_div(Rn, w), kRB);
    nblk = nb * w;
    if (c >= nblk) return;
    by = c / nb; b = c % nb;
  }
  __shared__ double sh[5 * 32];
  __shared__ int64_t shl[2][16];
  const int idx = b * kRB + threadIdx.x;
  double e = 0.0, m = 0.0;
  int64_t ei = 0, mi = 0;
  if (d.coding == kSSC) {
    if (idx < Rn) { e = d.IE[idx]; m = d.M[idx]; ei = d.IE[idx]; mi = d.M[idx]; }
  } else {
    const int rgn = ceil_div(Rn, w);
    if (idx < rgn && ((d.amask[buf][idx] >> by) & 1)) {
      const size_t k = (size_t)idx * w + by;
      e = d.IE[k]; m = d.M[k]; ei = d.IE[k]; mi = d.M[k];
    }
  }
  if (m < 0) m = -m;
  double v[5] = {e, e * e, m, m * m, 0.0};
  v[4] = v[3] * v[3];
  ref_block_sums<5>(v, sh);
  const int64_t se = block_sum_int<int64_t>(ei, shl[0]);
  const int64_t sm = block_sum_int<int64_t>(mi, shl[1]);
  if (threadIdx.x == 0) {
    double* o = d.avblk + (size_t)c * 5;
    for (int q = 0; q < 5; ++q) put_dev(o + q, v[q]);
    put_dev(d.iblk + 2 * c, se);
    put_dev(d.iblk + 2 * c + 1, sm);
  }
  // the state finish_step advances (R, buf, step, active) decides nothing above but active,
  // which retires late blocks as well
  if (last_block(&st->avdone, nblk)) finish_step(d, st, nblk);
}

// Spin-configuration hash of the new population (trace mode only): sum over 8-byte chunks
// i of mix64(mix64(i) ^ chunk_i), chunks numbered unit-major, zero-padded per unit.
__global__ __launch_bounds__(256) void k_hash(const RunDesc* __restrict__ runs,
                                              const int32_t* __restrict__ offs, int nruns) {
  const int* off = offs + kKUnits * (nruns + 1);
  const int r = find_run(off, nruns, blockIdx.x);
  const RunDesc& d = runs[r];
  RunState* st = d.st;
  if (!d.trace || !st->active || st->overflow) return;
  const int u = blockIdx.x - off[r];
  const int Rn = st->Rnew;
  const int units = (d.coding == kSSC) ? Rn : ceil_div(Rn, d.w);
  if (u >= units) return;
  const size_t bytes = (d.coding == kSSC) ? (size_t)d.N : (size_t)d.N * (d.w / 8);
  const size_t chunks = (bytes + 7) / 8;
  const uint8_t* p = (const uint8_t*)d.pop[st->buf ^ 1] + (size_t)u * bytes;
  uint64_t h = 0;
  for (size_t c = threadIdx.x; c < chunks; c += blockDim.x) {
    uint64_t v = 0;
    const size_t nbt = bytes - c * 8 < 8 ? bytes - c * 8 : 8;
    for (size_t i = 0; i < nbt; ++i) v |= (uint64_t)p[c * 8 + i] << (8 * i);
    h += mix64(mix64((uint64_t)u * chunks + c) ^ v);
  }
  __shared__ uint64_t shh[4];
  const uint64_t t = block_sum_int<uint64_t>(h, shh);
  if (threadIdx.x == 0)
    atomicAdd((unsigned long long*)&d.out[st->step].spin_hash, (unsigned long long)t);
}

// Step bookkeeping (SSC:620-632 device part): add the average partials in launch order,
// write the step record, advance R, IS (tau draws + sweep streams), Emean and the step.
// Run by the last k_avg block of the run; the partials pass through LDS 256 blocks at a time
// and thread 0 adds them.
__device__ void finish_step(const RunDesc& d, RunState* st, int nblk) {
  constexpr int kC = 256;
  __shared__ double fs[kC * 5];
  __shared__ int64_t fi[kC * 2];
  double sums[5] = {0, 0, 0, 0, 0};
  int64_t se = 0, sm = 0;
  for (int c0 = 0; c0 < nblk; c0 += kC) {
    const int m = min(kC, nblk - c0);
    for (int i = threadIdx.x; i < m * 5; i += blockDim.x) fs[i] = get_dev(d.avblk + (size_t)c0 * 5 + i);
    for (int i = threadIdx.x; i < m * 2; i += blockDim.x) fi[i] = get_dev(d.iblk + (size_t)c0 * 2 + i);
    __syncthreads();
    if (threadIdx.x == 0)
      for (int c = 0; c < m; ++c) {
        for (int q = 0; q < 5; ++q) sums[q] += fs[c * 5 + q];
        se += fi[2 * c];
        sm += fi[2 * c + 1];
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const int step = st->step, R = st->R, Rn = st->Rnew, w = d.w;
  StepOut& o = d.out[step];
  for (int q = 0; q < 5; ++q) o.sums[q] = sums[q];
  o.sumE = se;
  o.sumM = sm;
  o.R = Rn;
  o.Rprev = R;
  uint64_t IS = st->IS;
  if (d.coding == kSSC) IS += (uint64_t)R + (uint64_t)Rn * kStreamsPerUnit;
  else IS += (uint64_t)ceil_div(R, w) * w + (uint64_t)ceil_div(Rn, w) * kStreamsPerUnit;
  o.IS_after = IS;
  st->IS = IS;
  st->Emean = sums[0] / (double)Rn;
  st->R = Rn;
  st->buf ^= 1;
  st->step = step + 1;
  if (step + 1 > d.n_steps || step == st->pause) st->active = 0;
}

// Serving: re-activate the resident runs for a tick; run r executes steps up to pause[r].
__global__ void k_resume(const RunDesc* __restrict__ runs, int nruns,
                         const int32_t* __restrict__ pause) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nruns) return;
  RunState* st = runs[r].st;
  st->pause = pause[r];
  st->active = (!st->overflow && st->step <= runs[r].n_steps && st->step <= pause[r]) ? 1 : 0;
}

}  // namespace pa

// ============================================================================================
// Launch entry points used by the host engine (one per kernel, hipLaunchKernelGGL on the
// caller's stream).
// ============================================================================================
namespace pa {

template <bool INIT>
static hipError_t launch_sweep_msc(int w, int sw, int nt, int nblocks, size_t lds,
                                   hipStream_t s, const RunDesc* runs, const int32_t* offs,
                                   int nruns) {
#define PA_MSC(WT, ST, NTV)                                                                    \
  hipLaunchKernelGGL((k_sweep_msc<WT, ST, NTV, INIT>), dim3(nblocks), dim3(NTV), lds, s, runs, \
                     offs, nruns)
#define PA_MSC_NT(WT, ST)                   \
  do {                                      \
    if (nt == 1024) PA_MSC(WT, ST, 1024);   \
    else if (nt == 512) PA_MSC(WT, ST, 512); \
    else if (nt == 256) PA_MSC(WT, ST, 256); \
    else PA_MSC(WT, ST, 128);               \
  } while (0)
  if (w == 8) {
    PA_MSC_NT(uint8_t, uint8_t);
  } else if (w == 16) {
    if (sw == 16) PA_MSC_NT(uint16_t, uint16_t);
    else PA_MSC_NT(uint16_t, uint8_t);
  } else if (w == 32) {
    if (sw == 32) PA_MSC_NT(uint32_t, uint32_t);
    else if (sw == 16) PA_MSC_NT(uint32_t, uint16_t);
    else PA_MSC_NT(uint32_t, uint8_t);
  } else {
    if (sw == 64) PA_MSC_NT(uint64_t, uint64_t);
    else if (sw == 32) PA_MSC_NT(uint64_t, uint32_t);
    else if (sw == 16) PA_MSC_NT(uint64_t, uint16_t);
    else PA_MSC_NT(uint64_t, uint8_t);
  }
#undef PA_MSC_NT
#undef PA_MSC
  return hipGetLastError();
}

template <bool FAST>
static hipError_t launch_sweep_msc_bal(int w, int sw, int nt, int P, size_t lds, hipStream_t s,
                                       const RunDesc* runs, int nruns) {
#define PA_MSCB(WT, ST, NTV) \
  hipLaunchKernelGGL((k_sweep_msc_bal<WT, ST, NTV, FAST>), dim3(P), dim3(NTV), lds, s, runs, nruns, P)
#define PA_MSCB_NT(WT, ST)                     \
  do {                                         \
    if (nt == 512) PA_MSCB(WT, ST, 512);        \
    else if (nt == 256) PA_MSCB(WT, ST, 256);   \
    else PA_MSCB(WT, ST, 128);                  \
  } while (0)
  if (w < 32) return hipErrorInvalidValue;  // no balanced instance for 8- and 16-bit words
  if (w == 32) {
    if (sw == 32) PA_MSCB_NT(uint32_t, uint32_t);
    else if (sw == 16) PA_MSCB_NT(uint32_t, uint16_t);
    else PA_MSCB_NT(uint32_t, uint8_t);
  } else {
    if (sw == 64) PA_MSCB_NT(uint64_t, uint64_t);
    else if (sw == 32) PA_MSCB_NT(uint64_t, uint32_t);
    else if (sw == 16) PA_MSCB_NT(uint64_t, uint16_t);
    else PA_MSCB_NT(uint64_t, uint8_t);
  }
#undef PA_MSCB_NT
#undef PA_MSCB
  return hipGetLastError();
}

template <typename F>
static hipError_t allow_lds(F* f) {
  return hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(160 * 1024 - 2048));
}

// Streaming copy (16-byte vectors, grid-stride): reference point for HBM bandwidth and for
// scaling the FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                              size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace pa

extern "C" {

using namespace pa;

hipError_t pa_launch_resume(hipStream_t s, const RunDesc* runs, int nruns, const int32_t* pause) {
  if (nruns <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_resume, dim3((nruns + 255) / 256), dim3(256), 0, s, runs, nruns, pause);
  return hipGetLastError();
}

hipError_t pa_launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(256 * 64), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
  return hipGetLastError();
}

// Raise the dynamic-LDS limit of every kernel that keeps a lattice in LDS.
hipError_t pa_prepare_kernels(void) {
  hipError_t e = hipSuccess;
#define PA_ALLOW(...) if (e == hipSuccess) e = allow_lds(__VA_ARGS__)
  PA_ALLOW(k_init_ssc);
  PA_ALLOW((k_sweep_msc_half<uint32_t, false>));
  PA_ALLOW((k_sweep_msc_half<uint32_t, true>));
  PA_ALLOW((k_sweep_msc_half<uint64_t, false>));
  PA_ALLOW((k_sweep_msc_half<uint64_t, true>));
  PA_ALLOW((k_resample_msc<uint32_t, 256>));
  PA_ALLOW((k_resample_msc<uint32_t, 1024>));
  PA_ALLOW((k_resample_msc<uint64_t, 256>));
  PA_ALLOW((k_resample_win<uint32_t>));
  PA_ALLOW((k_resample_win<uint64_t>));
  PA_ALLOW((k_sweep_ssc<128, 0>));
  PA_ALLOW((k_sweep_ssc<256, 0>));
  PA_ALLOW((k_sweep_ssc<512, 0>));
  PA_ALLOW((k_sweep_ssc<128, 1>));
  PA_ALLOW((k_sweep_ssc<256, 1>));
  PA_ALLOW((k_sweep_ssc<512, 1>));
  PA_ALLOW((k_sweep_ssc<128, 2>));
  PA_ALLOW((k_sweep_ssc<128, 3>));
  PA_ALLOW((k_sweep_ssc<128, 4>));
  PA_ALLOW((k_sweep_ssc<256, 4>));
  PA_ALLOW((k_sweep_ssc<512, 4>));
  PA_ALLOW((k_sweep_ssc<128, 5>));
  PA_ALLOW((k_sweep_ssc<128, 6>));
  PA_ALLOW((k_sweep_ssc<256, 6>));
  PA_ALLOW((k_sweep_ssc<512, 6>));
  PA_ALLOW((k_sweep_ssc<256, 5>));
  PA_ALLOW((k_sweep_ssc<512, 5>));
  PA_ALLOW((k_sweep_ssc_q<1024, 2>));
  PA_ALLOW(k_sweep_ssc_p);
  PA_ALLOW(k_sweep_ssc_p2);
#define PA_ALLOW_MSC(WT, ST)                          \
  PA_ALLOW((k_sweep_msc_bal<WT, ST, 128, false>));    \
  PA_ALLOW((k_sweep_msc_bal<WT, ST, 256, false>));    \
  PA_ALLOW((k_sweep_msc_bal<WT, ST, 512, false>));    \
  PA_ALLOW((k_sweep_msc_bal<WT, ST, 128, true>));     \
  PA_ALLOW((k_sweep_msc_bal<WT, ST, 256, true>));     \
  PA_ALLOW((k_sweep_msc_bal<WT, ST, 512, true>));     \
  PA_ALLOW(k_sweep_msc<WT, ST, 128, false>);          \
  PA_ALLOW(k_sweep_msc<WT, ST, 256, false>);          \
  PA_ALLOW(k_sweep_msc<WT, ST, 512, false>);          \
  PA_ALLOW(k_sweep_msc<WT, ST, 128, true>);           \
  PA_ALLOW(k_sweep_msc<WT, ST, 256, true>);           \
  PA_ALLOW(k_sweep_msc<WT, ST, 512, true>);           \
  PA_ALLOW(k_sweep_msc<WT, ST, 1024, false>);         \
  PA_ALLOW(k_sweep_msc<WT, ST, 1024, true>)
  PA_ALLOW_MSC(uint32_t, uint32_t);
  PA_ALLOW_MSC(uint32_t, uint16_t);
  PA_ALLOW_MSC(uint32_t, uint8_t);
  PA_ALLOW_MSC(uint64_t, uint64_t);
  PA_ALLOW_MSC(uint64_t, uint32_t);
  PA_ALLOW_MSC(uint64_t, uint16_t);
  PA_ALLOW_MSC(uint64_t, uint8_t);
#undef PA_ALLOW_MSC
#define PA_ALLOW_MSC_NARROW(WT, ST)                   \
  PA_ALLOW(k_sweep_msc<WT, ST, 128, false>);          \
  PA_ALLOW(k_sweep_msc<WT, ST, 256, false>);          \
  PA_ALLOW(k_sweep_msc<WT, ST, 512, false>);          \
  PA_ALLOW(k_sweep_msc<WT, ST, 1024, false>);         \
  PA_ALLOW(k_sweep_msc<WT, ST, 128, true>);           \
  PA_ALLOW(k_sweep_msc<WT, ST, 256, true>);           \
  PA_ALLOW(k_sweep_msc<WT, ST, 512, true>);           \
  PA_ALLOW(k_sweep_msc<WT, ST, 1024, true>)
  PA_ALLOW_MSC_NARROW(uint8_t, uint8_t);
  PA_ALLOW_MSC_NARROW(uint16_t, uint16_t);
  PA_ALLOW_MSC_NARROW(uint16_t, uint8_t);
  PA_ALLOW((k_resample_msc<uint8_t, 256>));
  PA_ALLOW((k_resample_msc<uint16_t, 256>));
  PA_ALLOW((k_resample_win<uint8_t>));
  PA_ALLOW((k_resample_win<uint16_t>));
#undef PA_ALLOW_MSC_NARROW
#undef PA_ALLOW
  return e;
}

// Initial population of a group of runs sharing (coding, w, slice, threads).
hipError_t pa_launch_init(int coding, int w, int sw, int nt, int half, int nblocks, size_t lds,
                          hipStream_t s, const RunDesc* runs, const int32_t* offs, int nruns,
                          const uint64_t* is0) {
  if (nblocks <= 0) return hipSuccess;
  if (coding == kSSC) {
    hipLaunchKernelGGL(k_init_ssc, dim3(nblocks), dim3(256), lds, s, runs, offs, nruns, is0);
    return hipGetLastError();
  }
  if (w == 8)
    hipLaunchKernelGGL(k_init_msc<uint8_t>, dim3(nblocks), dim3(128), 0, s, runs, offs, nruns, is0);
  else if (w == 16)
    hipLaunchKernelGGL(k_init_msc<uint16_t>, dim3(nblocks), dim3(128), 0, s, runs, offs, nruns, is0);
  else if (w == 32)
    hipLaunchKernelGGL(k_init_msc<uint32_t>, dim3(nblocks), dim3(128), 0, s, runs, offs, nruns, is0);
  else
    hipLaunchKernelGGL(k_init_msc<uint64_t>, dim3(nblocks), dim3(128), 0, s, runs, offs, nruns, is0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (half) {
    if (w < 32) return hipErrorInvalidValue;  // one sublattice in LDS: 32-bit slices only
    if (w == 32)
      hipLaunchKernelGGL((k_sweep_msc_half<uint32_t, true>), dim3(nblocks), dim3(1024), lds, s, runs, offs, nruns);
    else
      hipLaunchKernelGGL((k_sweep_msc_half<uint64_t, true>), dim3(nblocks * 2), dim3(1024), lds, s, runs, offs, nruns);
    return hipGetLastError();
  }
  return launch_sweep_msc<true>(w, sw, nt, nblocks * (w / sw), lds, s, runs, offs, nruns);
}

hipError_t pa_occupancy(int kind, int coding, int w, int sw, int nt, size_t lds, int* blocks_per_cu) {
  if (kind != 11 && kind != 15) return hipErrorInvalidValue;
  if (w < 32) return hipErrorInvalidValue;  // no balanced instance for 8- and 16-bit words
  (void)coding;
#define PA_OCC(WT, ST, NTV)                                                                  \
  return kind == 15                                                                          \
             ? hipOccupancyMaxActiveBlocksPerMultiprocessor(                                 \
                   blocks_per_cu, k_sweep_msc_bal<WT, ST, NTV, true>, NTV, lds)              \
             : hipOccupancyMaxActiveBlocksPerMultiprocessor(                                 \
                   blocks_per_cu, k_sweep_msc_bal<WT, ST, NTV, false>, NTV, lds)
#define PA_OCC_NT(WT, ST) do { if (nt == 512) PA_OCC(WT, ST, 512); else if (nt == 256) PA_OCC(WT, ST, 256); else PA_OCC(WT, ST, 128); } while (0)
  if (w == 32) {
    if (sw == 32) PA_OCC_NT(uint32_t, uint32_t);
    else if (sw == 16) PA_OCC_NT(uint32_t, uint16_t);
    else PA_OCC_NT(uint32_t, uint8_t);
  } else {
    if (sw == 64) PA_OCC_NT(uint64_t, uint64_t);
    else if (sw == 32) PA_OCC_NT(uint64_t, uint32_t);
    else if (sw == 16) PA_OCC_NT(uint64_t, uint16_t);
    else PA_OCC_NT(uint64_t, uint8_t);
  }
#undef PA_OCC_NT
#undef PA_OCC
}

hipError_t pa_launch_step(int kind, int coding, int w, int sw, int nt, int nblocks, size_t lds,
                          hipStream_t s, const RunDesc* runs, const int32_t* offs, int nruns) {
  if (nblocks <= 0) return hipSuccess;
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_qpart, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns); break;
    case 1: hipLaunchKernelGGL(k_tau, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns); break;
    case 2: hipLaunchKernelGGL(k_scan, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns); break;
    case 3: hipLaunchKernelGGL(k_resample, dim3(nblocks), dim3(256), 0, s, runs, offs, nruns); break;
    case 4:
      if (coding == kSSC) {
        // for SSC the sw argument carries the sweep path (ssc_sweeps)
        if (sw == 2 || sw == 3) {
          if (nt != 128) return hipErrorInvalidValue;
          if (sw == 2) hipLaunchKernelGGL((k_sweep_ssc<128, 2>), dim3(nblocks), dim3(128), lds, s, runs, offs, nruns);
          else hipLaunchKernelGGL((k_sweep_ssc<128, 3>), dim3(nblocks), dim3(128), lds, s, runs, offs, nruns);
        } else if (sw == 4 || sw == 5) {
          if (sw == 4) {
            if (nt == 512) hipLaunchKernelGGL((k_sweep_ssc<512, 4>), dim3(nblocks), dim3(512), lds, s, runs, offs, nruns);
            else if (nt == 256) hipLaunchKernelGGL((k_sweep_ssc<256, 4>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
            else if (nt == 128) hipLaunchKernelGGL((k_sweep_ssc<128, 4>), dim3(nblocks), dim3(128), lds, s, runs, offs, nruns);
            else return hipErrorInvalidValue;
          } else {
            if (nt == 512) hipLaunchKernelGGL((k_sweep_ssc<512, 5>), dim3(nblocks), dim3(512), lds, s, runs, offs, nruns);
            else if (nt == 256) hipLaunchKernelGGL((k_sweep_ssc<256, 5>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
            else if (nt == 128) hipLaunchKernelGGL((k_sweep_ssc<128, 5>), dim3(nblocks), dim3(128), lds, s, runs, offs, nruns);
            else return hipErrorInvalidValue;
          }
        } else if (sw == 8 || sw == 9) {
          // packed-byte sweeps of any L % 8 == 0: RPW = nt / (32 P) replicas per workgroup
          // (9: the two-sweep step of L = 24)
          if (nt < 64 || nt > (sw == 8 ? 1024 : 512) || nt % 64) return hipErrorInvalidValue;
          if (sw == 8) hipLaunchKernelGGL(k_sweep_ssc_p, dim3(nblocks), dim3(nt), lds, s, runs, offs, nruns);
          else hipLaunchKernelGGL(k_sweep_ssc_p2, dim3(nblocks), dim3(nt), lds, s, runs, offs, nruns);
        } else if (sw == 7) {
          // packed-byte sweeps, two tasks per thread; nt is a whole number of waves
          if (nt < 64 || nt > 1024 || nt % 64) return hipErrorInvalidValue;
          hipLaunchKernelGGL((k_sweep_ssc_q<1024, 2>), dim3(nblocks), dim3(nt), lds, s, runs, offs, nruns);
        } else if (sw == 6) {
          if (nt == 512) hipLaunchKernelGGL((k_sweep_ssc<512, 6>), dim3(nblocks), dim3(512), lds, s, runs, offs, nruns);
          else if (nt == 256) hipLaunchKernelGGL((k_sweep_ssc<256, 6>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
          else hipLaunchKernelGGL((k_sweep_ssc<128, 6>), dim3(nblocks), dim3(128), lds, s, runs, offs, nruns);
        } else if (sw == 1) {
          if (nt == 512) hipLaunchKernelGGL((k_sweep_ssc<512, 1>), dim3(nblocks), dim3(512), lds, s, runs, offs, nruns);
          else if (nt == 256) hipLaunchKernelGGL((k_sweep_ssc<256, 1>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
          else hipLaunchKernelGGL((k_sweep_ssc<128, 1>), dim3(nblocks), dim3(128), lds, s, runs, offs, nruns);
        } else {
          if (nt == 512) hipLaunchKernelGGL((k_sweep_ssc<512, 0>), dim3(nblocks), dim3(512), lds, s, runs, offs, nruns);
          else if (nt == 256) hipLaunchKernelGGL((k_sweep_ssc<256, 0>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
          else hipLaunchKernelGGL((k_sweep_ssc<128, 0>), dim3(nblocks), dim3(128), lds, s, runs, offs, nruns);
        }
      } else {
        return launch_sweep_msc<false>(w, sw, nt, nblocks, lds, s, runs, offs, nruns);
      }
      break;
    case 5: hipLaunchKernelGGL(k_avg, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns); break;
    case 6: hipLaunchKernelGGL(k_hash, dim3(nblocks), dim3(256), 0, s, runs, offs, nruns); break;
    case 8: hipLaunchKernelGGL(k_qgroup, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns); break;
    case 11: return launch_sweep_msc_bal<false>(w, sw, nt, nblocks, lds, s, runs, nruns);
    case 15: return launch_sweep_msc_bal<true>(w, sw, nt, nblocks, lds, s, runs, nruns);
    case 16:  // MSC sweeps with one sublattice in LDS (32-bit slices, 1024 threads)
      if (w < 32) return hipErrorInvalidValue;
      if (w == 32)
        hipLaunchKernelGGL((k_sweep_msc_half<uint32_t, false>), dim3(nblocks), dim3(1024), lds, s, runs, offs, nruns);
      else
        hipLaunchKernelGGL((k_sweep_msc_half<uint64_t, false>), dim3(nblocks), dim3(1024), lds, s, runs, offs, nruns);
      break;
    case 12: hipLaunchKernelGGL(k_overlap, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns); break;
    case 13: hipLaunchKernelGGL(k_hist, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns, 0); break;
    case 14: hipLaunchKernelGGL(k_hist, dim3(nblocks), dim3(1024), 0, s, runs, offs, nruns, 1); break;
    case 10:  // nt: threads per workgroup (32-bit words: 256 or 1024)
      if (w == 8)
        hipLaunchKernelGGL((k_resample_msc<uint8_t, 256>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
      else if (w == 16)
        hipLaunchKernelGGL((k_resample_msc<uint16_t, 256>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
      else if (w == 32 && nt == 1024)
        hipLaunchKernelGGL((k_resample_msc<uint32_t, 1024>), dim3(nblocks), dim3(1024), lds, s, runs, offs, nruns);
      else if (w == 32)
        hipLaunchKernelGGL((k_resample_msc<uint32_t, 256>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
      else
        hipLaunchKernelGGL((k_resample_msc<uint64_t, 256>), dim3(nblocks), dim3(256), lds, s, runs, offs, nruns);
      break;
    case 17:  // MSC resampling by windows of new groups (whole 128-byte site chunks)
      if (w == 8)
        hipLaunchKernelGGL((k_resample_win<uint8_t>), dim3(nblocks), dim3(kWinJ), kWinLds, s, runs, offs, nruns);
      else if (w == 16)
        hipLaunchKernelGGL((k_resample_win<uint16_t>), dim3(nblocks), dim3(kWinJ), kWinLds, s, runs, offs, nruns);
      else if (w == 32)
        hipLaunchKernelGGL((k_resample_win<uint32_t>), dim3(nblocks), dim3(kWinJ), kWinLds, s, runs, offs, nruns);
      else
        hipLaunchKernelGGL((k_resample_win<uint64_t>), dim3(nblocks), dim3(kWinJ), kWinLds, s, runs, offs, nruns);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // extern "C"
