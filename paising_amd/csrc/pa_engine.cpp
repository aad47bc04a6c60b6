// Host engine: schedules, exact acceptance thresholds, device buffers, batched step launches
// (optionally replayed from a hipGraph), host thermodynamics and the C ABI of paising.h.
//
// Reference counterparts (PAisingSSC.cu = SSC, PAisingMSC.cu = MSC):
//   schedule B_i = B_{i-1} + dB while B_i <= Bfin      SSC:508-510,661 / MSC:670-672,812
//   Boltzmann tables                                   SSC:512-516     / MSC:674-675
//   thermodynamics E, C, M, M2, M4, BF, S              SSC:620-630     / MSC:772-782
//   results row                                        SSC:679-690     / MSC:829-840
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/paising.h"
#include "pa_dev.h"
#include "pa_prims.h"

extern "C" hipError_t pa_launch_init(int coding, int w, int sw, int nt, int half, int nblocks, size_t lds,
                                     hipStream_t s, const pa::RunDesc* runs, const int32_t* offs,
                                     int nruns, const uint64_t* is0);
extern "C" hipError_t pa_launch_step(int kind, int coding, int w, int sw, int nt, int nblocks,
                                     size_t lds, hipStream_t s, const pa::RunDesc* runs,
                                     const int32_t* offs, int nruns);
extern "C" hipError_t pa_prepare_kernels(void);
extern "C" hipError_t pa_occupancy(int kind, int coding, int w, int sw, int nt, size_t lds,
                                   int* blocks_per_cu);
extern "C" hipError_t pa_launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
extern "C" hipError_t pa_launch_resume(hipStream_t s, const pa::RunDesc* runs, int nruns,
                                       const int32_t* pause);

namespace {

using namespace pa;

constexpr int kNBmax = 10000;            // nBmax, SSC:57
constexpr size_t kLdsLimit = 163840;     // gfx950 LDS per workgroup

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHECK(expr)                                                                   \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(PA_ERR_HIP, std::string("GPUassert: ") + hipGetErrorString(e_) + " ; " \
                                  + __FILE__ + " ; line " + std::to_string(__LINE__));   \
  } while (0)

int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Reference schedule: FP64 accumulation of dB from Binit while <= Bfin.
std::vector<double> schedule(double Binit, double dB, double Bfin, int max_steps) {
  std::vector<double> B{Binit};
  double b = Binit + dB;
  while (b <= Bfin) {
    if (max_steps > 0 && (int)B.size() - 1 >= max_steps) break;
    B.push_back(b);
    if ((int)B.size() > kNBmax + 1) break;
    b = B.back() + dB;
  }
  return B;
}

// V(T) = #{v in [0, 2^32) : uniform_double(v) < T}; uniform_double is non-decreasing in v.
uint64_t ssc_threshold(double T, int umap) {
  uint64_t lo = 0, hi = 1ull << 32;  // find the smallest v with u(v) >= T (2^32 if none)
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (uniform_double((uint32_t)mid, umap) >= T) hi = mid; else lo = mid + 1;
  }
  return lo;
}

struct SweepVariant {
  int coding, w, sw, nt, path;  // path: SSC sweep kernel (1 = register-cached 3-block loop)
  bool operator<(const SweepVariant& o) const {
    return std::tie(coding, w, sw, nt, path) < std::tie(o.coding, o.w, o.sw, o.nt, o.path);
  }
};

// ---- memory of the server: grow-only pools, stream-ordered copies, deferred frees ----
// A plain run or batch allocates its device memory once, at creation, with hipMalloc.  The
// server admits and retires runs while earlier ticks are still on the device, so it must not
// allocate or free device memory per request (hipFree waits for the whole device).  It
// carves run arenas and batch tables out of large device slabs that only grow, stages host
// data in pinned slabs so copies stay asynchronous on the server stream, and returns a block
// to its pool only after an event recorded behind the last work that may touch it has
// completed.

// First-fit allocator over slabs of one kind of memory (device or pinned host); blocks are
// 256-byte aligned, free neighbours coalesce.
class SlabPool {
 public:
  explicit SlabPool(bool host) : host_(host) {}
  ~SlabPool() { release_all(); }
  SlabPool(const SlabPool&) = delete;
  SlabPool& operator=(const SlabPool&) = delete;

  // A free block of `bytes`, or null when no slab has room (the caller may grow()).
  void* take(size_t bytes) {
    bytes = round(bytes);
    for (Slab& s : slabs_) {
      for (auto it = s.free_.begin(); it != s.free_.end(); ++it) {
        if (it->second < bytes) continue;
        const size_t off = it->first, len = it->second;
        s.free_.erase(it);
        if (len > bytes) s.free_[off + bytes] = len - bytes;
        uint8_t* p = s.base + off;
        used_[p] = bytes;
        in_use_ += bytes;
        return p;
      }
    }
    return nullptr;
  }
  // Returns a block to its slab (the caller guarantees that no queued work touches it).
  void give(void* ptr) {
    auto u = used_.find((uint8_t*)ptr);
    if (u == used_.end()) return;
    const size_t bytes = u->second;
    used_.erase(u);
    in_use_ -= bytes;
    for (Slab& s : slabs_) {
      uint8_t* p = (uint8_t*)ptr;
      if (p < s.base || p >= s.base + s.size) continue;
      size_t off = (size_t)(p - s.base), len = bytes;
      auto next = s.free_.lower_bound(off);
      if (next != s.free_.end() && off + len == next->first) {
        len += next->second;
        next = s.free_.erase(next);
      }
      if (next != s.free_.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
          prev->second += len;
          return;
        }
      }
      s.free_[off] = len;
      return;
    }
  }
  // One more slab of at least `bytes`.
  hipError_t grow(size_t bytes) {
    bytes = round(bytes);
    void* p = nullptr;
    hipError_t e = host_ ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();  // the failure is reported to the caller, not left pending
      return e;
    }
    Slab s;
    s.base = (uint8_t*)p;
    s.size = bytes;
    s.free_[0] = bytes;
    slabs_.push_back(std::move(s));
    reserved_ += bytes;
    ++grows_;
    return hipSuccess;
  }
  // Frees the slabs without a block in use (only when the device is idle: hipFree waits).
  void trim() {
    for (size_t i = 0; i < slabs_.size();) {
      Slab& s = slabs_[i];
      if (s.free_.size() == 1 && s.free_.begin()->second == s.size) {
        free_mem(s.base);
        reserved_ -= s.size;
        slabs_.erase(slabs_.begin() + i);
      } else {
        ++i;
      }
    }
  }
  void release_all() {
    for (Slab& s : slabs_) free_mem(s.base);
    slabs_.clear();
    used_.clear();
    reserved_ = in_use_ = 0;
  }
  size_t reserved() const { return reserved_; }
  size_t in_use() const { return in_use_; }
  int grows() const { return grows_; }
  int slabs() const { return (int)slabs_.size(); }

 private:
  struct Slab {
    uint8_t* base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
  };
  static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; }
  void free_mem(void* p) {
    if (host_) (void)hipHostFree(p);
    else (void)hipFree(p);
  }
  bool host_;
  std::vector<Slab> slabs_;
  std::map<uint8_t*, size_t> used_;
  size_t reserved_ = 0, in_use_ = 0;
  int grows_ = 0;
};

// The server's memory: device and pinned pools on one stream.  Work is fenced once per tick:
// blocks handed back during a tick (arenas of retired runs, tables of a replaced batch,
// staging copies, replaced graphs) are released when that tick's fence event completes.
// Fences complete in order (one stream), so completion is a sequence number.
class ServerMem {
 public:
  ServerMem() : dev(false), pin(true) {}
  ~ServerMem() {
    for (Fence& f : fences_) (void)hipEventDestroy(f.ev);
    for (hipEvent_t e : spare_) (void)hipEventDestroy(e);
    for (Fence& f : fences_) release(f.items);
    release(pending_);
  }
  hipStream_t stream = nullptr;
  size_t dev_cap = SIZE_MAX;          // bound on device slab bytes (the HBM budget + tables)
  static constexpr size_t kDevSlab = (size_t)4 << 30;
  static constexpr size_t kPinSlab = (size_t)64 << 20;
  static constexpr size_t kTabSlab = (size_t)16 << 20;
  SlabPool dev, pin;
  SlabPool tab{false};  // batch tables (KBs per run): slabs of 16 MiB outside dev_cap

  // Device block of `bytes`, growing the pool by a slab of max(bytes, 4 GiB) within dev_cap.
  // Returns 0 with the block in *out; 1 when dev_cap leaves no room now (blocks come back as
  // runs retire); 2 on a HIP error (*err).
  int alloc_dev(size_t bytes, void** out, hipError_t* err) {
    *out = dev.take(bytes);
    if (*out) return 0;
    const size_t need = bytes + 256;
    if (dev.reserved() + need > dev_cap) {
      if (dev.in_use() != 0) return 1;
      dev.trim();  // nothing in use, no work queued on the pool: give back the empty slabs
      if (dev.reserved() + need > dev_cap) return 1;
    }
    const size_t slab = std::max(need, std::min(kDevSlab, dev_cap - dev.reserved()));
    *err = dev.grow(slab);
    if (*err != hipSuccess) return 2;
    *out = dev.take(bytes);
    if (!*out) { *err = hipErrorOutOfMemory; return 2; }
    return 0;
  }
  hipError_t alloc_tab(size_t bytes, void** out) {
    *out = tab.take(bytes);
    if (*out) return hipSuccess;
    hipError_t e = tab.grow(std::max(bytes + 256, kTabSlab));
    if (e != hipSuccess) return e;
    *out = tab.take(bytes);
    return *out ? hipSuccess : hipErrorOutOfMemory;
  }
  void* alloc_pin(size_t bytes) {
    void* p = pin.take(bytes);
    if (!p && pin.grow(std::max(bytes + 256, kPinSlab)) == hipSuccess) p = pin.take(bytes);
    return p;
  }
  // Stream-ordered copy of host data (staged in pinned memory, released after the fence).
  hipError_t upload(void* dst, const void* src, size_t bytes) {
    if (!bytes) return hipSuccess;
    void* st = alloc_pin(bytes);
    if (!st) return hipErrorOutOfMemory;
    memcpy(st, src, bytes);
    pending_.push_back(Item{Item::kPin, st, nullptr, nullptr});
    return hipMemcpyAsync(dst, st, bytes, hipMemcpyHostToDevice, stream);
  }
  void free_dev_later(void* p) { if (p) pending_.push_back(Item{Item::kDev, p, nullptr, nullptr}); }
  void free_tab_later(void* p) { if (p) pending_.push_back(Item{Item::kTab, p, nullptr, nullptr}); }
  void free_pin_later(void* p) { if (p) pending_.push_back(Item{Item::kPin, p, nullptr, nullptr}); }
  void destroy_graph_later(hipGraphExec_t x, hipGraph_t g) {
    if (x || g) pending_.push_back(Item{Item::kGraph, nullptr, x, g});
  }
  // Records the fence of the work enqueued so far; returns its sequence number.
  hipError_t fence(int64_t* seq) {
    hipEvent_t e;
    if (!spare_.empty()) { e = spare_.back(); spare_.pop_back(); }
    else {
      hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
      if (r != hipSuccess) return r;
    }
    hipError_t r = hipEventRecord(e, stream);
    if (r != hipSuccess) { spare_.push_back(e); return r; }
    Fence f;
    f.ev = e;
    f.seq = ++issued_;
    f.items.swap(pending_);
    fences_.push_back(std::move(f));
    if (seq) *seq = issued_;
    return hipSuccess;
  }
  // Releases what completed fences held (no wait); returns the last completed sequence.
  int64_t poll() {
    while (!fences_.empty()) {
      hipError_t q = hipEventQuery(fences_.front().ev);
      if (q == hipErrorNotReady) break;
      retire_front();
    }
    return done_;
  }
  // Waits for the oldest outstanding fence (used only when the server has nothing to enqueue).
  hipError_t wait_oldest() {
    if (fences_.empty()) return hipSuccess;
    hipError_t e = hipEventSynchronize(fences_.front().ev);
    if (e == hipSuccess) retire_front();
    return e;
  }
  int64_t done() const { return done_; }
  int64_t issued() const { return issued_; }
  bool idle() const { return fences_.empty(); }
  bool has_pending() const { return !pending_.empty(); }

 private:
  struct Item {
    enum Kind { kDev, kTab, kPin, kGraph } kind;
    void* p;
    hipGraphExec_t gx;
    hipGraph_t g;
  };
  struct Fence {
    hipEvent_t ev = nullptr;
    int64_t seq = 0;
    std::vector<Item> items;
  };
  void release(std::vector<Item>& items) {
    for (Item& it : items) {
      if (it.kind == Item::kDev) dev.give(it.p);
      else if (it.kind == Item::kTab) tab.give(it.p);
      else if (it.kind == Item::kPin) pin.give(it.p);
      else {
        if (it.gx) (void)hipGraphExecDestroy(it.gx);
        if (it.g) (void)hipGraphDestroy(it.g);
      }
    }
    items.clear();
  }
  void retire_front() {
    Fence& f = fences_.front();
    release(f.items);
    done_ = f.seq;
    spare_.push_back(f.ev);
    fences_.pop_front();
  }
  std::deque<Fence> fences_;
  std::vector<Item> pending_;
  std::vector<hipEvent_t> spare_;
  int64_t issued_ = 0, done_ = 0;
};

struct RunHost {
  pa_params p{};
  int w = 1, N = 0, n_steps = 0, cap_pop = 0, cap_units = 0;
  int R0 = 0;
  int sw = 0, nt = 0, path = 0;  // MSC path 1: one sublattice in LDS (msc_unit_half)
  int ssc_p = 1, rpw = 1;  // SSC path 8: parts per stream quad, replicas per workgroup
  bool half = false;
  int rs_sites = 1;
  int rs_split = 1;
  int rs_lp = 1;
  bool rs_win = false;  // MSC resampling by windows of new groups (k_resample_win)
  std::vector<double> B;
  // adaptive schedule: B grows as steps are decided; deltaBeta carries over (SSC:645-660)
  bool adaptive = false;
  bool adaptive_done = false;
  double deltaBeta = 0.0;
  size_t footprint = 0;
  // device memory (one allocation carved into pieces); a served run's arena belongs to the
  // server's pool, which takes it back behind the run's last queued work
  void* arena = nullptr;
  bool pooled = false;
  RunDesc desc{};
  std::vector<StepOut> out_h;
  RunState st_h{};
  bool results_valid = false;

  ~RunHost() {
    if (arena && !pooled) (void)hipFree(arena);
  }
  int blocks(int kind) const {
    switch (kind) {
      case kKUnits: return cap_units;
      case kKSweep: return ceil_div(cap_units, rpw);
      case kKChunks: return ceil_div(cap_pop, kRedBlock);
      case kKGroupChunks: return ceil_div(cap_units, kRedBlock);
      case kKAvg: return p.coding == kSSC ? ceil_div(cap_pop, kRedBlock)
                                          : ceil_div(cap_units, kRedBlock) * w;
      // multiples of 8 blocks per run keep every run's first block on XCD 0 (the dispatcher
      // deals workgroups to the 8 XCDs round-robin); k_resample_msc relies on it
      case kKResample:
        if (rs_win) return ceil_div(cap_units, kWinJ) * rs_split;
        return p.coding == kSSC ? ceil_div(cap_units, 8) * 8
                                : ceil_div(ceil_div(N, rs_sites), 8 * rs_lp) * 8 * rs_lp * rs_split;
    }
    return 0;
  }
  // LDS of a packed SSC sweep workgroup (path 8): RPW lattices, and for one part per stream
  // quad the geometry table (16 bytes per quad and dword)
  size_t sweep_lds() const {
    return (size_t)rpw * N + (ssc_p == 1 ? (size_t)32 * 16 * ceil_div(N / 2, 128) : 0);
  }
  // LDS of an MSC sweep unit: its lattice slice (one sublattice for the half-resident form),
  // plus for L = 32 over 512 threads (four parts, one word of the stream's block each) the
  // Philox words handed between the parts: 2 x 4 half sweeps x 4 words x 128 streams
  size_t msc_unit_lds() const {
    const size_t slice = (size_t)(half ? N / 2 : N) * (sw / 8);
    return slice + (!half && N / 2 == 512 && nt == 512 ? (size_t)2 * 4 * 4 * 128 * 4 : 0);
  }
  size_t resample_lds() const {
    // parents' chunks plus one zero chunk
    if (rs_win) return kWinLds;
    return p.coding == kSSC ? 0 : (size_t)(cap_units + 1) * rs_sites * (w / 8);
  }
};

// Autotuner cache: shape -> best (slice_bits, threads, balance) measured by pa_autotune.
struct TuneKey {
  int coding, w, L, rbucket;
  bool operator<(const TuneKey& o) const {
    return std::tie(coding, w, L, rbucket) < std::tie(o.coding, o.w, o.L, o.rbucket);
  }
};
struct TuneVal {
  int slice_bits, threads, balance;
  double ms;
};
std::map<TuneKey, TuneVal>& tune_cache() {
  static std::map<TuneKey, TuneVal> c;
  return c;
}
TuneKey tune_key(const pa_params& p) {
  int rb = 0;
  while (((int64_t)1 << (rb + 1)) <= (int64_t)p.Rinit) ++rb;  // floor(log2 Rinit)
  return TuneKey{p.coding, p.coding == PA_MSC ? p.msbits : 1, p.L, rb};
}

// Default MSC slice width.  A sweep unit is one (group, slice) pair and runs on one CU; its
// cost is about N (27 + 5 sw) VALU ops per sweep (27 per word, 5 per bit).  Narrow slices
// give more units but repeat the per-word work, so the width minimises
// max(one unit, all units / CUs): full words when the launch already fills the chip, narrow
// slices for a few groups.  groups counts every group of the launch in units of this
// lattice's sites.
int device_cus();
int default_slice(int w, int N, int64_t groups, bool half_ok, bool* half) {
  const double cus = device_cus();
  int best = 0;
  double tbest = 0;
  *half = false;
  for (int sw = w; sw >= 8; sw /= 2) {
    // the slice must fit LDS, or (32-bit slices) one sublattice of it with the other in
    // registers
    const bool full = (size_t)N * (sw / 8) + 2048 <= kLdsLimit;
    const bool hf = !full && half_ok && sw == 32;
    if (!full && !hf) continue;
    const double c = 27.0 + 5.0 * sw;
    const double t = std::max(c, (double)groups * (w / sw) * c / cus);
    if (!best || t < tbest) { best = sw; tbest = t; *half = hf; }
  }
  return best ? best : 8;
}

// Validate, size and compute the per-step tables of a run (no device work).  co_work: sweep
// work of the other runs that share this run's launches (MSC groups x sites, single-spin
// replicas x sites), 0 for a lone run.
int plan_run(const pa_params& p_in, RunHost& h, double cap_sigma, int64_t co_work = 0) {
  pa_params p = p_in;
  {
    auto it = tune_cache().find(tune_key(p));
    if (it != tune_cache().end()) {
      if (!p.slice_bits) p.slice_bits = it->second.slice_bits;
      if (!p.threads) p.threads = it->second.threads;
      if (!p.balance) p.balance = it->second.balance;
    }
  }
  if (p.L < 4 || (p.L & 1) || p.L > 1024) return fail(PA_ERR_INVALID, "L must be even, 4..1024");
  if (p.coding != PA_SSC && p.coding != PA_MSC) return fail(PA_ERR_INVALID, "coding");
  if (p.coding == PA_MSC && p.msbits != 8 && p.msbits != 16 && p.msbits != 32 && p.msbits != 64)
    return fail(PA_ERR_INVALID, "msbits must be 8, 16, 32 or 64");
  if (p.Rinit < 1) return fail(PA_ERR_INVALID, "Rinit < 1");
  if (!(p.dB > 0.0) || p.Binit < 0.0) return fail(PA_ERR_INVALID, "dB must be > 0, Binit >= 0");
  if (p.uniform != PA_UNIFORM_CUDA101 && p.uniform != PA_UNIFORM_CUDA100)
    return fail(PA_ERR_INVALID, "uniform");
  if (p.schedule_len < 0 || (p.schedule_len > 0 && !p.schedule))
    return fail(PA_ERR_INVALID, "schedule_len > 0 needs a schedule");
  if (p.adaptive && p.schedule_len > 0)
    return fail(PA_ERR_INVALID, "adaptive and an explicit schedule exclude each other");
  if (p.adaptive) {
    if (p.min_overlap == 0.0 && p.max_overlap == 0.0) {
      p.min_overlap = 0.85;  // MinOverlap, SSC:41
      p.max_overlap = 0.87;  // MaxOverlap, SSC:42
    }
    if (!(p.min_overlap > 0.0 && p.min_overlap < p.max_overlap && p.max_overlap <= 1.0))
      return fail(PA_ERR_INVALID, "adaptive step needs 0 < min_overlap < max_overlap <= 1");
  }
  p.schedule = nullptr;  // copied below; the caller's array is not kept
  h.p = p;
  h.w = p.coding == PA_SSC ? 1 : p.msbits;
  h.N = p.L * p.L;
  if (p_in.schedule_len > 0) {
    // replay of an explicit schedule (later runs of an adaptive -M series, SSC:662)
    h.B.assign(p_in.schedule, p_in.schedule + p_in.schedule_len);
    if (p.max_steps > 0 && (int)h.B.size() > p.max_steps + 1) h.B.resize(p.max_steps + 1);
    h.n_steps = (int)h.B.size() - 1;
  } else if (p.adaptive) {
    // steps are decided one at a time after each step; the tables hold up to nBmax steps
    h.adaptive = true;
    h.deltaBeta = p.dB;
    h.B = {p.Binit};
    if (p.Binit + p.dB <= p.Bfin) h.B.push_back(p.Binit + p.dB);
    h.adaptive_done = h.B.size() == 1;
    h.n_steps = h.adaptive_done ? 0 : (p.max_steps > 0 ? std::min(p.max_steps, kNBmax - 1) : kNBmax - 1);
  } else {
    h.B = schedule(p.Binit, p.dB, p.Bfin, p.max_steps);
    h.n_steps = (int)h.B.size() - 1;
  }
  if (h.n_steps + 1 > kNBmax || (!h.adaptive && h.n_steps + 1 >= kNBmax))
    return fail(PA_ERR_NBMAX, "Error: number of temperature steps exceeds nBmax=10000.");
  h.R0 = p.Rinit;
  const double margin = cap_sigma * std::sqrt((double)p.Rinit) + 64.0;
  if ((double)p.Rinit + margin + 64.0 >= 2147483647.0)
    return fail(PA_ERR_INVALID, "Rinit too large: population capacity must stay below 2^31");
  h.cap_pop = p.Rinit + (int)margin;
  if (p.coding == PA_SSC) {
    h.cap_units = h.cap_pop;
  } else {
    h.cap_units = ceil_div(h.cap_pop, h.w);
    h.cap_pop = h.cap_units * h.w;
  }
  // sweep variant
  const int Nh = h.N / 2;
  const bool aligned = (Nh % 512) == 0;
  // aligned lattices: 128 streams x P parts, P = blocks per stream per half sweep, <= 4
  const int nb = Nh / 512;
  // unaligned lattices (MSC; SSC beyond 9 sites per stream): the Philox blocks of a stream's
  // half sweep are dealt to P threads
  const int nblk_u = ceil_div(ceil_div(Nh, 128), 4) + 1;
  const int P_u = Nh <= 512 ? 1 : nblk_u >= 8 ? 4 : nblk_u >= 4 ? 2 : 1;
  // packed-byte SSC sweeps of any lattice with L % 8 == 0 (k_sweep_ssc_p, path 8) unless
  // the aligned packed kernel takes it (decided below) or slice_bits = 1 asks for the
  // site-per-thread kernels: one part per stream quad (the whole stream walk, RPW = NT / 32
  // replicas per workgroup, no workgroup barrier) up to 12 dwords per quad and half sweep;
  // beyond that P parts deal the Philox blocks of a half sweep, P the power of two nearest
  // to a third of the blocks (measured at R = 20000 against P = 2 .. 32: L = 56 P = 2, 112
  // P = 4, 144 P = 8, 208 and 240 P = 16 were the fastest).  Its word counter of a step must
  // stay within 32 bits.
  const int nq8 = ceil_div(Nh, 128);
  int P8 = 1;
  if (nq8 > 12) {
    const double third = (ceil_div(nq8, 4) + 1) / 3.0;
    P8 = 2;
    while (P8 < 32 && third > P8 * 1.41421356) P8 *= 2;
  }
  // One part gives a replica 32 lanes (a quarter of the site kernels' waves), so with at
  // most 4 dwords per quad it needs enough replicas in the launch (this run's plus, in a
  // batch, co_work / N of the other single-spin runs): from 5000 with up to 3 dwords, 16000
  // with 4 (lone runs, sweep attempts/ns, packed against site kernels: L = 16 R = 4000 790
  // vs 825, 6000 998 vs 886; L = 32 R = 8000 1661 vs 1769, 16000 1922 vs 1945, 20000 1933
  // vs 1779; L = 40, 48 from R = 2000 on 695 vs 663 and 870 vs 802).
  const int64_t reps8 = (int64_t)p.Rinit + co_work / h.N;
  const bool few8 = P8 == 1 && ((nq8 <= 3 && reps8 < 5000) || (nq8 == 4 && reps8 < 16000));
  const bool packed8 = p.coding == PA_SSC && p.L % 8 == 0 && p.slice_bits != 1 &&
                       (!aligned || p.L < 64 || p.L > 256 || p.slice_bits == 2) &&
                       (!few8 || p.slice_bits == 2) &&
                       2.0 * p.eqsweeps * nq8 < 4294967295.0;
  h.nt = p.threads ? p.threads : 128 * (aligned ? std::min(4, std::max(1, nb)) : P_u);
  if (packed8) {
    // a given workgroup size caps the parts at threads / 32
    const int nt8 = p.threads ? p.threads : P8 == 1 ? 256 : 64 * ceil_div(32 * P8, 64);
    if (nt8 % 64 || nt8 < 64 || nt8 > 1024)
      return fail(PA_ERR_INVALID, "threads must be a multiple of 64 from 64 to 1024 for "
                                  "packed-byte SSC sweeps");
    h.ssc_p = std::min(P8, nt8 / 32);
    h.rpw = nt8 / (32 * h.ssc_p);
    if (h.sweep_lds() + 1024 > kLdsLimit)
      return fail(PA_ERR_INVALID, "packed-byte SSC workgroup exceeds LDS");
    h.nt = nt8;
    // one part and 257..384 sites per sublattice (L = 24): the two-sweep step with the lanes
    // sorted by dword count (k_sweep_ssc_p2, at most 512 threads, whole wavefronts per class)
    h.path = h.ssc_p == 1 && nt8 <= 512 && Nh > 256 && Nh <= 384 &&
                     (h.rpw * ((Nh - 256) / 4)) % 64 == 0
                 ? 9
                 : 8;
  } else if (h.nt != 128 && h.nt != 256 && h.nt != 512 && !(h.nt == 1024 && p.coding == PA_MSC)) {
    // 1024-thread workgroups (MSC only): twice the waves per CU for lattices whose slice
    // leaves room for one workgroup per CU
    return fail(PA_ERR_INVALID, "threads must be 128, 256 or 512 (or 1024 for MSC)");
  }
  if (!aligned && p.coding == PA_SSC && ceil_div(Nh, 128) <= 9 && !packed8) h.nt = 128;
  // MSC lattices of at most 9 sites per stream and half sweep (L <= 48): 256-thread
  // workgroups (R = 20000, attempts/ns against 128 threads: L = 16 2091 vs 1976, 24 1953
  // vs 1860, 32 3370 vs 3159, 40 2347 vs 1915)
  if (p.coding == PA_MSC && !p.threads && Nh <= 9 * 128 && h.nt < 256) h.nt = 256;
  // SSC sweep kernel path (pa_kernels.hip, ssc_sweeps): columns repeating every third site
  // of a stream (L = 96, 192) for more than one Philox block per thread and half sweep (6),
  // register-cached block loop for two or three blocks per thread (1), fixed-column block
  // loop for more blocks when L/2 divides 128 (4), register-cached stream walk for unaligned
  // lattices with at most 4 / 9 sites per stream and half sweep (2 / 3), blocks dealt to
  // several threads per stream beyond that (5)
  if (p.coding == PA_SSC && !packed8) {
    const int ntmax = ceil_div(Nh, 128);
    const int Ld2 = p.L / 2;
    // (only while Ld2 < 128: a stream's first site is then its only one in row 0, and its
    // last its only one in row L - 1; L = 384 takes the general block loop)
    const bool period3 = Ld2 < 128 && 128 % Ld2 != 0 && 384 % Ld2 == 0 && (384 / Ld2) % 2 == 0;
    if (aligned && period3 && nb > h.nt / 128) h.path = 6;
    else if (aligned && nb > h.nt / 128 && nb <= 3 * (h.nt / 128)) h.path = 1;
    else if (aligned && nb > 3 * (h.nt / 128) && 128 % (p.L / 2) == 0) h.path = 4;
    else if (!aligned && ntmax <= 4) h.path = 2;
    else if (!aligned && ntmax <= 9) h.path = 3;
    else if (!aligned) h.path = 5;
    // packed-byte sweeps (k_sweep_ssc_q) for aligned lattices from L = 64 to 256 unless a
    // block shape is forced: one task per (four streams, Philox block), T = 32 nb tasks per
    // half sweep, two per thread, the workgroup the power of two >= T / 2 (sweep kernel alone,
    // R = 20000, attempts/ns, against the previous paths: L = 64 2151-2332 vs 1931, 96 1879-1940
    // vs 1829, 128 2480 vs 1974, 160 2014 vs 1019, 224 1927, 256 2004 vs 1712).  L = 192 keeps
    // its period-3 path (2006 vs 1831: 1152 tasks leave 44 % of the 1024-thread slots idle),
    // L = 32 its own (one task per stream quad leaves most of a wave idle).  A replica then
    // has only T / 128 waves, so a small population takes the other paths until its waves
    // reach eight per CU (L = 64 R = 1000: 984 packed against 1479; R = 20000: 2233 vs 1909).
    if (aligned && p.L >= 64 && p.L <= 256 && !(period3 && nb > 16) && !p.threads &&
        p.slice_bits != 1) {
      const int T = 32 * nb;
      int nt = 64;
      while (nt < 1024 && 2 * nt < T) nt *= 2;
      if ((int64_t)h.cap_units * (nt / 64) >= 8 * (int64_t)device_cus()) {
        h.path = 7;
        h.nt = nt;
      }
    }
  }
  if (p.coding == PA_MSC) {
    // 32-bit slices with one sublattice in LDS and the other in registers (msc_unit_half):
    // aligned lattices with L/2 dividing 128, at most 8 Philox blocks per thread of 1024
    const bool half_ok = h.w >= 32 && aligned && 128 % (p.L / 2) == 0 && nb <= 64 &&
                         (size_t)Nh * 4 + 2048 <= kLdsLimit && (!p.threads || p.threads == 1024) &&
                         p.balance <= 0;
    int sw = p.slice_bits;
    bool half = false;
    if (sw == 0) sw = default_slice(h.w, h.N, ceil_div(p.Rinit, h.w) + co_work / h.N, half_ok, &half);
    if (sw != 8 && sw != 16 && sw != 32 && sw != 64) return fail(PA_ERR_INVALID, "slice_bits");
    if (sw > h.w) sw = h.w;
    if ((size_t)h.N * (sw / 8) + 1024 > kLdsLimit) {
      if (!(sw == 32 && half_ok)) return fail(PA_ERR_INVALID, "lattice slice exceeds LDS");
      half = true;
    }
    h.sw = sw;
    h.half = half;
    // 5 to 13 sites per stream and half sweep (L = 40, 48, 56; words of 32 bits or more):
    // 512-thread workgroups where the population fills the balanced sweep, 1024 (eight
    // parts per stream, more waves per unit) where it has fewer units than CUs (R = 20000 /
    // 5000, attempts/ns, against the 256-thread default: L = 40 3034 vs 2774, 1704 vs 1221;
    // 48 3390 vs 3027, 2113 vs 1471; 56 3358 vs 3065)
    // L = 32 (one Philox block per stream and half sweep), lone runs: 512 threads, the four
    // parts of a stream taking one word of its block each (R = 5000 / 20000 / 80000,
    // attempts/ns, against 256 threads with one working part: 2059 vs 1270, 4091 vs 3326,
    // 5427 vs 4710).  In a batch the other runs bring the waves, and a variant of its own
    // would add a launch per step (the small-run mix: 662 against 693-723 runs/s).
    if (!p.threads && !half && Nh == 512 && co_work == 0) h.nt = 512;
    if (!p.threads && h.w >= 32 && Nh > 4 * 128 && Nh <= 13 * 128) {
      const int64_t units = ((int64_t)ceil_div(p.Rinit, h.w) + co_work / h.N) * (h.w / sw);
      h.nt = units * 9 / 10 >= device_cus() ? 512 : 1024;
    }
    if (half) {
      h.nt = 1024;
      h.path = 1;
    }
    // a slice over 80 KiB leaves room for one workgroup per CU: 1024 threads keep 4 waves
    // per SIMD busy where 512 would leave 2
    if (!p.threads && (aligned ? nb >= 8 : nblk_u >= 16) && (size_t)h.N * (sw / 8) > 80 * 1024)
      h.nt = 1024;
    // resampling stages a 64-byte chunk of every parent group in LDS (<= 96 KiB); the
    // chunks of one 128-byte line are resampled on one XCD so the line is fetched once
    h.rs_sites = 64 / (h.w / 8);
    while (h.rs_sites > 1 && (size_t)h.cap_units * h.rs_sites * (h.w / 8) > 96 * 1024) h.rs_sites /= 2;
    h.rs_lp = 128 / (h.rs_sites * (h.w / 8));
    // populations whose 64-byte chunks of every parent group exceed the LDS budget are
    // resampled by windows of new groups instead (whole 128-byte chunks: 32 / 16 words)
    const int win_sites = 128 / (h.w / 8);
    h.rs_win = h.rs_sites < 64 / (h.w / 8) && h.N % win_sites == 0;
    if (h.rs_win) {
      // windows x site-chunk ranges: about eight workgroups per CU in all
      h.rs_sites = win_sites;
      h.rs_split = std::max(1, std::min(h.N / win_sites,
                                        ceil_div(8 * device_cus(), ceil_div(h.cap_units, kWinJ))));
    } else if ((size_t)h.cap_units * h.rs_sites * (h.w / 8) > 150 * 1024) {
      return fail(PA_ERR_INVALID, "population too large for one device resampling chunk");
    }
    // splits of a site chunk share its new groups (1024 threads per workgroup for 32-bit
    // populations above 256 groups, 256 otherwise) and run on one XCD, but each stages the
    // chunk of every parent group again: at most about one new group per thread, and only
    // while the chunks alone leave the device short of workgroups (a 256x256 lattice has
    // 32768 chunks of 16 bytes; 13 splits of each were 22 GB of staging per step)
    if (!h.rs_win)
      h.rs_split = std::max(1, std::min({16, ceil_div(h.cap_units, h.w == 32 ? 1024 : 256),
                                         ceil_div(8 * device_cus(), ceil_div(h.N, h.rs_sites))}));
  } else {
    if ((size_t)h.N + 1024 > kLdsLimit) return fail(PA_ERR_INVALID, "lattice exceeds LDS");
    h.sw = 8;
  }
  return PA_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte layout of a run's arena; returns total size.  Pointers are filled when base != null.
size_t layout(RunHost& h, uint8_t* base) {
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = align_up(o + bytes); return base ? base + at : nullptr; };
  const size_t unit_bytes = h.p.coding == PA_SSC ? (size_t)h.N : (size_t)h.N * (h.w / 8);
  RunDesc& d = h.desc;
  d.pop[0] = take((size_t)h.cap_units * unit_bytes);
  d.pop[1] = take((size_t)h.cap_units * unit_bytes);
  d.amask[0] = (uint64_t*)take((size_t)h.cap_units * 8);
  d.amask[1] = (uint64_t*)take((size_t)h.cap_units * 8);
  d.IE = (int32_t*)take((size_t)h.cap_pop * 4);
  d.M = (int32_t*)take((size_t)h.cap_pop * 4);
  d.Roff = (uint32_t*)take((size_t)h.cap_pop * 4);
  d.pre = (uint32_t*)take((size_t)h.cap_pop * 4);
  d.par = (int32_t*)take((size_t)h.cap_pop * 4);
  d.flags = (uint32_t*)take((size_t)h.cap_units * 8 * 4);
  d.gsum = (double*)take((size_t)h.cap_units * 8);
  d.hist = h.p.mhr ? (int32_t*)take((size_t)(h.N + 1) * 4) : nullptr;
  d.cblk = (uint32_t*)take((size_t)h.blocks(kKChunks) * 4);
  d.qblk = (double*)take((size_t)h.blocks(kKGroupChunks) * 8 + 8 * (size_t)h.blocks(kKChunks));
  d.avblk = (double*)take((size_t)h.blocks(kKAvg) * 5 * 8);
  d.iblk = (int64_t*)take((size_t)h.blocks(kKAvg) * 2 * 8);
  d.hblk = (uint64_t*)take(64);
  d.beta = (const double*)take((size_t)(h.n_steps + 1) * 8);
  d.dbeta = (const double*)take((size_t)(h.n_steps + 1) * 8);
  d.thr = (const uint64_t*)take((size_t)(h.n_steps + 1) * 3 * 8);
  d.out = (StepOut*)take((size_t)(h.n_steps + 1) * sizeof(StepOut));
  d.st = (RunState*)take(sizeof(RunState));
  return o;
}

// Acceptance thresholds of one step (3 words per step in the device table).
void step_thresholds(const RunHost& h, double b, uint64_t* t) {
  if (h.p.coding == PA_SSC) {
    // T[idxT0 + 2 idxT1] = min(1, exp(-2 (2 idxT0 - 1)(2 idxT1 - 4) B)), SSC:512-516:
    // dE <= 0 -> 1, dE = 4 -> exp(-4B), dE = 8 -> exp(-8B)
    t[0] = ssc_threshold(1.0, h.p.uniform);
    t[1] = ssc_threshold(std::min(1.0, std::exp(-4 * b)), h.p.uniform);
    t[2] = ssc_threshold(std::min(1.0, std::exp(-8 * b)), h.p.uniform);
  } else {
    t[0] = (uint32_t)std::ceil(4294967296. * std::exp(-4 * b));  // MSC:674
    t[1] = (uint32_t)std::ceil(4294967296. * std::exp(-8 * b));  // MSC:675
    t[2] = 0;
  }
}

// The scalar fields of a planned run's descriptor (geometry, coding, schedule length,
// stream seed, resampling shape); pointers are set by the caller.
void desc_scalars(RunHost& h) {
  RunDesc& d = h.desc;
  d.L = h.p.L; d.Ld2 = h.p.L / 2; d.N = h.N; d.Nh = h.N / 2;
  d.coding = h.p.coding; d.w = h.w; d.Rinit = h.p.Rinit; d.n_steps = h.n_steps;
  d.cap_units = h.cap_units; d.umap = h.p.uniform; d.eqsweeps = h.p.eqsweeps;
  d.trace = h.p.trace; d.q128 = 128 / d.Ld2; d.r128 = 128 % d.Ld2; d.seed = h.p.seed;
  d.rs_sites = h.rs_sites;
  d.rs_lp = h.rs_lp;
  d.rs_split = h.rs_split;
  d.ld2_magic = (uint32_t)((1ull << 32) / (uint64_t)d.Ld2 + 1);
  d.ssc_p = h.ssc_p;
}

constexpr int kWaitForMemory = -1;  // internal: the server pool has no room for now

// Allocate and upload a planned run's state (tables, initial RunState); no kernels.  With a
// server pool the arena comes from the pool and the uploads are stream-ordered copies on the
// server stream (kWaitForMemory: no room in the pool until runs retire).
int materialize(RunHost& h, ServerMem* mem = nullptr) {
  h.footprint = layout(h, nullptr);
  if (mem) {
    hipError_t err = hipSuccess;
    const int r = mem->alloc_dev(h.footprint, &h.arena, &err);
    if (r == 1) return kWaitForMemory;
    if (r == 2) return fail(PA_ERR_HIP, std::string("GPUassert: ") + hipGetErrorString(err) +
                                            " (server pool, " + std::to_string(h.footprint) + " bytes)");
    h.pooled = true;
  } else {
    HIPCHECK(hipMalloc(&h.arena, h.footprint));
  }
  layout(h, (uint8_t*)h.arena);
  desc_scalars(h);
  auto put = [&](void* dst, const void* src, size_t n) {
    return mem ? mem->upload(dst, src, n) : hipMemcpy(dst, src, n, hipMemcpyHostToDevice);
  };
  auto zero = [&](void* dst, size_t n) {
    return mem ? hipMemsetAsync(dst, 0, n, mem->stream) : hipMemset(dst, 0, n);
  };
  RunDesc& d = h.desc;
  std::vector<double> beta(h.n_steps + 1, 0.0), dbeta(h.n_steps + 1, 0.0);
  std::vector<uint64_t> thr((size_t)(h.n_steps + 1) * 3, 0);
  beta[0] = h.B[0];
  const int known = std::min(h.n_steps, (int)h.B.size() - 1);  // adaptive: step 1 only
  for (int i = 1; i <= known; ++i) {
    beta[i] = h.B[i];
    dbeta[i] = h.B[i] - h.B[i - 1];
    step_thresholds(h, h.B[i], &thr[3 * i]);
  }
  HIPCHECK(put((void*)d.beta, beta.data(), (h.n_steps + 1) * 8));
  HIPCHECK(put((void*)d.dbeta, dbeta.data(), (h.n_steps + 1) * 8));
  HIPCHECK(put((void*)d.thr, thr.data(), thr.size() * 8));
  HIPCHECK(zero(d.out, (h.n_steps + 1) * sizeof(StepOut)));
  HIPCHECK(zero(d.flags, (size_t)h.cap_units * 8 * 4));
  if (d.hist) HIPCHECK(zero(d.hist, (size_t)(h.N + 1) * 4));
  RunState st{};
  st.step = 1;
  st.R = h.p.Rinit;
  st.Rnew = h.p.Rinit;
  st.buf = 0;
  const int units0 = h.p.coding == PA_SSC ? h.p.Rinit : ceil_div(h.p.Rinit, h.w);
  st.IS = h.p.initial_sequence + (uint64_t)units0 * kStreamsPerUnit;  // SSC:486 / MSC:647
  st.Emean = 0.0;                                                     // SSC:483
  st.lnQ = 0.0;
  st.overflow = 0;
  st.err = 0;
  st.active = h.n_steps > 0 ? 1 : 0;
  HIPCHECK(put(d.st, &st, sizeof(RunState)));
  h.st_h = st;
  return PA_OK;
}

// Kernel attributes (dynamic LDS limits) set once per process before the first launch.
int prepare_kernels_once() {
  static bool prepared = false;
  if (!prepared) {
    HIPCHECK(pa_prepare_kernels());
    prepared = true;
  }
  return PA_OK;
}

struct Group {  // runs sharing one sweep kernel variant
  SweepVariant v;
  std::vector<int> members;
  RunDesc* desc_d = nullptr;
  int32_t* offs_d = nullptr;
  int units_blocks = 0;  // total kKUnits blocks
  size_t lds = 0;
  int sweep_blocks = 0;  // SSC path 8: total kKSweep blocks and their LDS (RPW lattices)
  size_t sweep_lds = 0;
  int bal_workers = 0;   // > 0: makespan-balanced MSC sweep with this many workers
  size_t bal_lds = 0;
  int bal_kind = 11;     // 15: every member takes the register-cached path (fewer registers)
};

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
    else
      cus = 256;
  }
  return cus;
}

// Balanced sweeps pay off when a launch holds only a few units per CU (whole units per CU
// would leave CUs with n and n+1 units).  k workers per CU are made co-resident by padding
// their LDS so that no more than k fit a CU; P = k * CUs stays below the expected unit
// count so that a share always exceeds one unit (a unit is cut at most once).
void plan_balance(Group& g, const std::vector<RunHost*>& runs) {
  g.bal_workers = 0;
  // no balanced 1024-thread kernel, and none for 8- and 16-bit words
  if (g.v.coding != PA_MSC || g.v.nt > 512 || g.v.w < 32) return;
  int64_t units = 0;
  int force = 0;
  bool fast = true;
  for (int m : g.members) {
    const RunHost& h = *runs[m];
    if (h.p.eqsweeps == 0) return;
    units += (int64_t)h.cap_units * (h.w / h.sw);
    force = h.p.balance ? h.p.balance : force;
    const int Nh = h.N / 2;
    fast = fast && Nh % 512 == 0 && Nh / 512 <= g.v.nt / 128 && 128 % (h.p.L / 2) == 0;
  }
  if (force < 0) return;
  // not by default for lattices of at most 4 sites per stream and half sweep (L <= 32):
  // a balanced worker keeps one wave per SIMD busy, where whole units per workgroup fill
  // more (R = 20000: L = 16 2091 vs 1611, 24 1953 vs 1489, 32 3370 vs 2577 attempts/ns)
  for (int m : g.members)
    if (force == 0 && runs[m]->N / 2 <= 4 * 128) return;
  const int cus = device_cus();
  const int64_t expect = units * 9 / 10;
  if (force == 0 && (expect < cus || expect >= 12 * (int64_t)cus)) return;
  // the largest k <= expect / CUs (at most 4) for which exactly k workers fit a CU
  auto workers = [&](int kind, size_t* lds_out) {
    int k = (int)std::max<int64_t>(1, std::min<int64_t>(expect / cus, 4));
    for (; k >= 1; --k) {
      const size_t lds = std::max(g.lds, (size_t)(160 * 1024) / k - 4096);
      int occ = 0;
      if (pa_occupancy(kind, g.v.coding, g.v.w, g.v.sw, g.v.nt, lds, &occ) != hipSuccess) return 0;
      if (occ == k) { *lds_out = lds; return k; }
    }
    return 0;
  };
  // the instance without the other paths (fewer registers) only where it fits more workers
  // per CU: at equal worker counts the general instance runs about 1 % faster
  size_t lds11 = 0, lds15 = 0;
  const int k11 = workers(11, &lds11);
  const int k15 = fast ? workers(15, &lds15) : 0;
  g.bal_kind = k15 > k11 ? 15 : 11;
  const int k = std::max(k11, k15);
  if (k < 1) return;
  g.bal_lds = g.bal_kind == 15 ? lds15 : lds11;
  g.bal_workers = k * cus;
}

struct Batch {
  std::vector<RunHost*> runs;  // not owned when borrowed by the server
  std::vector<std::unique_ptr<RunHost>> owned;
  hipStream_t stream = nullptr;
  RunDesc* desc_d = nullptr;
  int32_t* offs_d = nullptr;
  uint64_t* is0_d = nullptr;
  int32_t* pause_d = nullptr;  // serving: last step of the tick per run
  int blocks[kKindCount] = {0, 0, 0, 0, 0};
  // MSC word widths 8, 16, 32, 64 -> index 0..3
  static constexpr int kWidths = 4;
  static int width_index(int w) { return w == 8 ? 0 : w == 16 ? 1 : w == 32 ? 2 : 3; }
  static int width_of(int wi) { return 8 << wi; }
  size_t rs_lds[kWidths] = {0, 0, 0, 0};  // MSC resampling LDS per word width
  int rs_nt[kWidths] = {256, 256, 256, 256};  // MSC resampling threads per workgroup
  bool has_msc = false;
  std::vector<Group> groups;
  static constexpr int kRsKinds = 1 + 2 * kWidths;
  struct RsGroup {  // the runs of one resampling kernel (SSC; per width plain and windowed)
    RunDesc* desc_d = nullptr;
    int32_t* offs_d = nullptr;
    int m = 0, blocks = 0;
  } rs_groups[kRsKinds];
  bool any_trace = false;
  bool any_mhr = false;
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int steps_launched = 0;
  int max_steps = 0;
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // timed steps: three events each (before resampling, before sweeps, after sweeps),
  // recorded on the run's stream and read back by collect_timing() after a sync
  std::vector<hipEvent_t> tev;
  size_t tev_used = 0;
  double sweep_ms = 0.0, resample_ms = 0.0, total_ms = 0.0;
  int timed_steps = 0;
  std::vector<void*> table_allocs;
  // serving: tables come from the server's pool and are released behind its fences, the
  // uploads are stream-ordered; null for plain runs and batches (hipMalloc at creation)
  ServerMem* mem = nullptr;
  bool no_side = false;  // a batch that only launches initial populations: no side streams

  ~Batch() { release(); }
  void release_tables() {
    if (mem) {
      mem->destroy_graph_later(gexec, graph);
      gexec = nullptr;
      graph = nullptr;
      for (void* p : table_allocs) mem->free_tab_later(p);
    } else {
      if (gexec) { (void)hipGraphExecDestroy(gexec); gexec = nullptr; }
      if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
      for (void* p : table_allocs) (void)hipFree(p);
    }
    table_allocs.clear();
    groups.clear();
    for (RsGroup& g : rs_groups) g = RsGroup{};
    init_is0.clear();
  }
  void release() {
    release_tables();
    if (ev0) { (void)hipEventDestroy(ev0); (void)hipEventDestroy(ev1); ev0 = ev1 = nullptr; }
    for (hipEvent_t e : tev) (void)hipEventDestroy(e);
    tev.clear();
    tev_used = 0;
    for (hipStream_t x : side) (void)hipStreamDestroy(x);
    for (hipEvent_t e : join_ev) (void)hipEventDestroy(e);
    side.clear();
    join_ev.clear();
    if (fork_ev) { (void)hipEventDestroy(fork_ev); fork_ev = nullptr; }
    if (stream) { (void)hipStreamDestroy(stream); stream = nullptr; }
  }

  int alloc_table(void** p, size_t bytes) {
    if (mem) HIPCHECK(mem->alloc_tab(bytes, p));
    else HIPCHECK(hipMalloc(p, bytes));
    table_allocs.push_back(*p);
    return PA_OK;
  }
  hipError_t put(void* dst, const void* src, size_t bytes) {
    return mem ? mem->upload(dst, src, bytes) : hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
  }

  // Build descriptor and prefix tables for the current run set.
  int build_tables() {
    release_tables();
    const int n = (int)runs.size();
    if (n == 0) return PA_OK;
    std::vector<RunDesc> descs(n);
    std::vector<int32_t> offs((size_t)kKindCount * (n + 1), 0);
    std::vector<uint64_t> is0(n);
    any_trace = false;
    any_mhr = false;
    max_steps = 0;
    for (int k = 0; k < kKindCount; ++k) {
      int acc = 0;
      for (int r = 0; r < n; ++r) {
        offs[(size_t)k * (n + 1) + r] = acc;
        acc += runs[r]->blocks(k);
      }
      offs[(size_t)k * (n + 1) + n] = acc;
      blocks[k] = acc;
    }
    for (int wi = 0; wi < kWidths; ++wi) {
      rs_lds[wi] = 0;
      rs_nt[wi] = 256;
    }
    has_msc = false;
    for (int r = 0; r < n; ++r) {
      descs[r] = runs[r]->desc;
      is0[r] = runs[r]->p.initial_sequence;
      if (runs[r]->p.coding == PA_MSC) has_msc = true;
      if (runs[r]->p.coding == PA_MSC && !runs[r]->rs_win) {
        const int wi = width_index(runs[r]->w);
        rs_lds[wi] = std::max(rs_lds[wi], runs[r]->resample_lds());
        // 1024-thread resampling workgroups pay off for 32-bit populations above 256 groups
        if (runs[r]->w == 32 && runs[r]->cap_units > 256) rs_nt[wi] = 1024;
      }
      any_trace |= runs[r]->p.trace != 0;
      any_mhr |= runs[r]->p.mhr != 0;
      max_steps = std::max(max_steps, runs[r]->n_steps);
    }
    int rc;
    if ((rc = alloc_table((void**)&desc_d, n * sizeof(RunDesc)))) return rc;
    if ((rc = alloc_table((void**)&offs_d, offs.size() * 4))) return rc;
    if ((rc = alloc_table((void**)&is0_d, n * 8))) return rc;
    if ((rc = alloc_table((void**)&pause_d, n * 4))) return rc;
    HIPCHECK(put(desc_d, descs.data(), n * sizeof(RunDesc)));
    HIPCHECK(put(offs_d, offs.data(), offs.size() * 4));
    HIPCHECK(put(is0_d, is0.data(), n * 8));
    // sweep groups (by kernel variant); the init launches use the same grouping
    {
      std::map<SweepVariant, std::vector<int>> gm;
      for (int r = 0; r < n; ++r) {
        const RunHost& h = *runs[r];
        SweepVariant v{h.p.coding, h.w, h.sw, h.nt, h.path};
        gm[v].push_back(r);
      }
      for (auto& kv : gm) {
        Group g;
        g.v = kv.first;
        g.members = kv.second;
        const int m = (int)g.members.size();
        std::vector<RunDesc> gd(m);
        std::vector<int32_t> go((size_t)kKindCount * (m + 1), 0);
        std::vector<uint64_t> gis(m);
        int acc = 0, acc8 = 0;
        size_t lds = 0, lds8 = 0;
        for (int i = 0; i < m; ++i) {
          const RunHost& h = *runs[g.members[i]];
          gd[i] = h.desc;
          gis[i] = h.p.initial_sequence;
          go[(size_t)kKUnits * (m + 1) + i] = acc;
          go[(size_t)kKSweep * (m + 1) + i] = acc8;
          acc += h.cap_units;
          acc8 += h.blocks(kKSweep);
          const size_t l = h.p.coding == PA_SSC ? (size_t)h.N
                           : h.msc_unit_lds();
          lds = std::max(lds, l);
          lds8 = std::max(lds8, h.path == 8 || h.path == 9 ? h.sweep_lds() : 0);
        }
        go[(size_t)kKUnits * (m + 1) + m] = acc;
        go[(size_t)kKSweep * (m + 1) + m] = acc8;
        g.units_blocks = acc;
        g.lds = lds;
        g.sweep_blocks = acc8;
        g.sweep_lds = lds8;
        if ((rc = alloc_table((void**)&g.desc_d, m * sizeof(RunDesc)))) return rc;
        if ((rc = alloc_table((void**)&g.offs_d, go.size() * 4))) return rc;
        HIPCHECK(put(g.desc_d, gd.data(), m * sizeof(RunDesc)));
        HIPCHECK(put(g.offs_d, go.data(), go.size() * 4));
        uint64_t* gis_d = nullptr;
        if ((rc = alloc_table((void**)&gis_d, m * 8))) return rc;
        HIPCHECK(put(gis_d, gis.data(), m * 8));
        plan_balance(g, runs);
        groups.push_back(g);
        init_is0.push_back(gis_d);
      }
    }
    // resampling tables per kernel (SSC; MSC per word width, plain and windowed): each
    // launch covers only the runs its kernel serves, not the blocks of the other runs
    auto rs_kind = [](const RunHost& h) {
      return h.p.coding == PA_SSC ? 0 : 1 + width_index(h.w) + (h.rs_win ? kWidths : 0);
    };
    for (int k = 0; k < kRsKinds; ++k) {
      RsGroup& g = rs_groups[k];
      g = RsGroup{};
      std::vector<RunDesc> gd;
      for (int r = 0; r < n; ++r)
        if (rs_kind(*runs[r]) == k) gd.push_back(runs[r]->desc);
      g.m = (int)gd.size();
      if (!g.m) continue;
      std::vector<int32_t> go((size_t)kKindCount * (g.m + 1), 0);
      int acc = 0, i = 0;
      for (int r = 0; r < n; ++r) {
        const RunHost& h = *runs[r];
        if (rs_kind(h) != k) continue;
        go[(size_t)kKResample * (g.m + 1) + i++] = acc;
        acc += h.blocks(kKResample);
      }
      go[(size_t)kKResample * (g.m + 1) + g.m] = acc;
      g.blocks = acc;
      if ((rc = alloc_table((void**)&g.desc_d, g.m * sizeof(RunDesc)))) return rc;
      if ((rc = alloc_table((void**)&g.offs_d, go.size() * 4))) return rc;
      HIPCHECK(put(g.desc_d, gd.data(), g.m * sizeof(RunDesc)));
      HIPCHECK(put(g.offs_d, go.data(), go.size() * 4));
    }
    return no_side ? PA_OK : ensure_side_streams();
  }
  std::vector<uint64_t*> init_is0;

  bool has_ssc() const {
    for (const RunHost* h : runs)
      if (h->p.coding == PA_SSC) return true;
    return false;
  }

  int launch_init() {
    for (size_t i = 0; i < groups.size(); ++i) {
      const Group& g = groups[i];
      HIPCHECK(pa_launch_init(g.v.coding, g.v.w, g.v.sw, g.v.nt,
                              g.v.coding == PA_MSC && g.v.path == 1, g.units_blocks, g.lds, stream,
                              g.desc_d, g.offs_d, (int)g.members.size(), init_is0[i]));
    }
    // MHR: the initial population enters the histogram too (SSC:500)
    if (any_mhr)
      HIPCHECK(pa_launch_step(14, 0, 0, 0, 0, blocks[kKChunks], 0, stream, desc_d, offs_d,
                              (int)runs.size()));
    return PA_OK;
  }

  // One beta step of every active run (5 launches, 6 with single-spin runs, + one sweep launch
  // per variant).
  int timing_event(hipEvent_t* e) {
    if (tev_used == tev.size()) {
      hipEvent_t x;
      HIPCHECK(hipEventCreate(&x));
      tev.push_back(x);
    }
    *e = tev[tev_used++];
    return PA_OK;
  }

  // Adds the recorded per-step kernel times to sweep_ms / resample_ms; call after a sync.
  int collect_timing() {
    for (size_t i = 0; i + 2 < tev_used; i += 3) {
      float a = 0, b = 0;
      HIPCHECK(hipEventElapsedTime(&a, tev[i], tev[i + 1]));
      HIPCHECK(hipEventElapsedTime(&b, tev[i + 1], tev[i + 2]));
      resample_ms += a;
      sweep_ms += b;
      ++timed_steps;
    }
    tev_used = 0;
    return PA_OK;
  }

  // The parts of a step in launch order; the stage entry points launch them one at a time.
  // q partials, lnQ and copy counts (QKer, CalcTauKer).
  int launch_counts() {
    const int n = (int)runs.size();
    if (has_msc)
      HIPCHECK(pa_launch_step(8, 0, 0, 0, 0, blocks[kKChunks], 0, stream, desc_d, offs_d, n));
    if (has_ssc())
      HIPCHECK(pa_launch_step(0, 0, 0, 0, 0, blocks[kKGroupChunks], 0, stream, desc_d, offs_d, n));
    HIPCHECK(pa_launch_step(1, 0, 0, 0, 0, blocks[kKChunks], 0, stream, desc_d, offs_d, n));
    return PA_OK;
  }
  // exclusive prefix of the copy counts and the parent of every child (CalcParSum)
  int launch_scan() {
    const int n = (int)runs.size();
    HIPCHECK(pa_launch_step(2, 0, 0, 0, 0, blocks[kKChunks], 0, stream, desc_d, offs_d, n));
    return PA_OK;
  }
  int launch_resample() {
    const RsGroup& gs = rs_groups[0];
    if (gs.m) HIPCHECK(pa_launch_step(3, 0, 0, 0, 0, gs.blocks, 0, stream, gs.desc_d, gs.offs_d, gs.m));
    for (int wi = 0; wi < kWidths; ++wi) {
      const RsGroup& g = rs_groups[1 + wi];
      if (g.m)
        HIPCHECK(pa_launch_step(10, PA_MSC, width_of(wi), 0, rs_nt[wi], g.blocks, rs_lds[wi], stream,
                                g.desc_d, g.offs_d, g.m));
      const RsGroup& gw = rs_groups[1 + kWidths + wi];
      if (gw.m)
        HIPCHECK(pa_launch_step(17, PA_MSC, width_of(wi), 0, kWinJ, gw.blocks, kWinLds, stream,
                                gw.desc_d, gw.offs_d, gw.m));
    }
    return PA_OK;
  }
  // EqSw sweeps and energies, one launch per kernel variant.  With several variants the
  // launches after the first go to side streams forked from the batch stream and joined back
  // into it (parallel branches of a captured step), so a variant's tail and the small
  // variants overlap the others instead of queueing behind them.  The variants' runs are
  // disjoint.  A balanced launch whose workers start late only waits longer: each waits on a
  // worker of higher index, and workgroups of one launch are dispatched in index order.
  int launch_group(const Group& g, hipStream_t s) {
    const int ns = g.v.coding == PA_SSC ? 1 : g.v.w / g.v.sw;
    if (g.v.coding == PA_MSC && g.v.path == 1)
      HIPCHECK(pa_launch_step(16, g.v.coding, g.v.w, 32, 1024, g.units_blocks * (g.v.w / 32),
                              g.lds, s, g.desc_d, g.offs_d, (int)g.members.size()));
    else if (g.bal_workers > 0)
      HIPCHECK(pa_launch_step(g.bal_kind, g.v.coding, g.v.w, g.v.sw, g.v.nt, g.bal_workers, g.bal_lds, s,
                              g.desc_d, g.offs_d, (int)g.members.size()));
    else if (g.v.coding == PA_SSC && (g.v.path == 8 || g.v.path == 9))
      HIPCHECK(pa_launch_step(4, g.v.coding, g.v.w, g.v.path, g.v.nt, g.sweep_blocks, g.sweep_lds, s,
                              g.desc_d, g.offs_d, (int)g.members.size()));
    else
      HIPCHECK(pa_launch_step(4, g.v.coding, g.v.w, g.v.coding == PA_SSC ? g.v.path : g.v.sw, g.v.nt,
                              g.units_blocks * ns, g.lds, s, g.desc_d, g.offs_d, (int)g.members.size()));
    return PA_OK;
  }
  int launch_sweeps() {
    const size_t ng = groups.size();
    int rc;
    if (ng < 2 || side.size() + 1 < ng) {
      for (const Group& g : groups)
        if ((rc = launch_group(g, stream))) return rc;
      return PA_OK;
    }
    HIPCHECK(hipEventRecord(fork_ev, stream));
    for (size_t i = 1; i < ng; ++i) {
      HIPCHECK(hipStreamWaitEvent(side[i - 1], fork_ev, 0));
      if ((rc = launch_group(groups[i], side[i - 1]))) return rc;
      HIPCHECK(hipEventRecord(join_ev[i - 1], side[i - 1]));
    }
    if ((rc = launch_group(groups[0], stream))) return rc;
    for (size_t i = 1; i < ng; ++i) HIPCHECK(hipStreamWaitEvent(stream, join_ev[i - 1], 0));
    return PA_OK;
  }
  // side streams and fork/join events for the sweep variants (outside any capture)
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> join_ev;
  hipEvent_t fork_ev = nullptr;
  int ensure_side_streams() {
    const size_t need = groups.size() > 1 ? groups.size() - 1 : 0;
    if (need && !fork_ev) HIPCHECK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    while (side.size() < need) {
      hipStream_t x;
      HIPCHECK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
      side.push_back(x);
      hipEvent_t e;
      HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      join_ev.push_back(e);
    }
    return PA_OK;
  }
  // optional hashes and histograms, then the averages (CalcAverages), whose last block per
  // run writes the step record and advances the run
  int launch_finish() {
    const int n = (int)runs.size();
    if (any_trace) HIPCHECK(pa_launch_step(6, 0, 0, 0, 0, blocks[kKUnits], 0, stream, desc_d, offs_d, n));
    if (any_mhr) HIPCHECK(pa_launch_step(13, 0, 0, 0, 0, blocks[kKChunks], 0, stream, desc_d, offs_d, n));
    HIPCHECK(pa_launch_step(5, 0, 0, 0, 0, blocks[kKAvg], 0, stream, desc_d, offs_d, n));
    return PA_OK;
  }

  int launch_step(bool time_sweep) {
    hipEvent_t t0 = nullptr, t1 = nullptr, t2 = nullptr;
    int rc;
    if (time_sweep)
      if ((rc = timing_event(&t0)) || (rc = timing_event(&t1)) || (rc = timing_event(&t2))) return rc;
    if ((rc = launch_counts()) || (rc = launch_scan())) return rc;
    if (time_sweep) HIPCHECK(hipEventRecord(t0, stream));
    if ((rc = launch_resample())) return rc;
    if (time_sweep) HIPCHECK(hipEventRecord(t1, stream));
    if ((rc = launch_sweeps())) return rc;
    if (time_sweep) HIPCHECK(hipEventRecord(t2, stream));
    return launch_finish();
  }

  int ensure_graph() {
    if (gexec) return PA_OK;
    HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_step(false);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &gr);
    if (rc) return rc;
    HIPCHECK(e);
    graph = gr;
    HIPCHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    return PA_OK;
  }

  // Adaptive schedule (plain single-run batches): decide B_i from the population after
  // step i-1 by the reference's growth + bisection search on the histogram overlap
  // (SSC:644-660, MSC:795-811), each overlap evaluated on the device (k_qgroup, k_qpart,
  // k_overlap; CalcOverlap SSC:317-334) with the block sums added on the host in launch
  // order; then write the step's beta, dB and thresholds.  Sets h.adaptive_done when the
  // schedule has passed Bfin.
  double stage_db = 0.0;
  std::vector<double> stage_part;
  std::vector<uint64_t> stage_thr;
  double stage_beta[2] = {0.0, 0.0};
  int adaptive_next(RunHost& h, int i) {
    if (i >= kNBmax)
      return fail(PA_ERR_NBMAX,
                  "Error: number of temperature steps exceeds nBmax=10000.\n Please consider "
                  "increasing the population size or decreasing the value of MinOverlap or "
                  "increasing the value of nBmax.");
    HIPCHECK(hipStreamSynchronize(stream));
    RunState st;
    HIPCHECK(hipMemcpy(&st, h.desc.st, sizeof st, hipMemcpyDeviceToHost));
    if (st.overflow) return fail(PA_ERR_CAPACITY, "population outgrew its buffers");
    if (st.err) return fail(PA_ERR_HIP, "device error bits " + std::to_string(st.err));
    if (st.step != i) return fail(PA_ERR_STATE, "adaptive step out of sequence");
    const int n = (int)runs.size();
    const int R = st.R, w = h.w;
    const int nblk = h.p.coding == PA_SSC ? ceil_div(R, kRedBlock)
                                          : ceil_div(ceil_div(R, w), kRedBlock) * w;
    stage_part.resize((size_t)nblk * 5);
    int rc = PA_OK;
    auto overlap = [&](double dB) -> double {
      stage_db = dB;
      hipError_t e = hipMemcpyAsync((void*)(h.desc.dbeta + i), &stage_db, 8, hipMemcpyHostToDevice, stream);
      if (e == hipSuccess && h.p.coding == PA_MSC)
        e = pa_launch_step(8, 0, 0, 0, 0, blocks[kKChunks], 0, stream, desc_d, offs_d, n);
      if (e == hipSuccess && h.p.coding == PA_SSC)
        e = pa_launch_step(0, 0, 0, 0, 0, blocks[kKGroupChunks], 0, stream, desc_d, offs_d, n);
      if (e == hipSuccess) e = pa_launch_step(12, 0, 0, 0, 0, blocks[kKAvg], 0, stream, desc_d, offs_d, n);
      if (e == hipSuccess)
        e = hipMemcpyAsync(stage_part.data(), h.desc.avblk, stage_part.size() * 8, hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) { rc = fail(PA_ERR_HIP, hipGetErrorString(e)); return 0.0; }
      double acc = 0.0;
      for (int c = 0; c < nblk; ++c) acc += stage_part[(size_t)c * 5];
      return acc / R;
    };
    const double Bprev = h.B[i - 1], Bfin = h.p.Bfin;
    const double MinOverlap = h.p.min_overlap, MaxOverlap = h.p.max_overlap;
    double ovl, dBmin = 0, dBmax = h.deltaBeta, dBmean;
    for (int it = 0;; ++it) {
      if (it > 100000) return fail(PA_ERR_STATE, "adaptive step: growth did not stop");
      ovl = overlap(dBmax);
      if (rc) return rc;
      if ((ovl >= MaxOverlap) && (Bprev + dBmax < Bfin)) dBmax *= 1.1;
      else break;
    }
    if (ovl >= MinOverlap) {
      dBmean = dBmax;
    } else {
      for (int it = 0;; ++it) {
        if (it > 100000) return fail(PA_ERR_STATE, "adaptive step: bisection did not converge");
        dBmean = 0.5 * (dBmin + dBmax);
        ovl = overlap(dBmean);
        if (rc) return rc;
        if (ovl < MinOverlap) dBmax = dBmean;
        else if (ovl >= MaxOverlap) dBmin = dBmean;
        else break;
      }
    }
    if ((Bprev < Bfin) && (Bprev + dBmean > Bfin)) h.deltaBeta = Bfin - Bprev;
    else h.deltaBeta = dBmean;
    const double Bi = Bprev + h.deltaBeta;
    if (!(Bi <= Bfin)) {
      h.adaptive_done = true;
      return PA_OK;
    }
    h.B.push_back(Bi);
    stage_beta[0] = Bi;
    stage_beta[1] = Bi - Bprev;
    stage_thr.assign(3, 0);
    step_thresholds(h, Bi, stage_thr.data());
    HIPCHECK(hipMemcpyAsync((void*)(h.desc.beta + i), &stage_beta[0], 8, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync((void*)(h.desc.dbeta + i), &stage_beta[1], 8, hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync((void*)(h.desc.thr + 3 * i), stage_thr.data(), 24, hipMemcpyHostToDevice, stream));
    return PA_OK;
  }

  int run_steps(int k) {
    for (int i = 0; i < k; ++i) {
      if (use_graph) {
        int rc = ensure_graph();
        if (rc) return rc;
        HIPCHECK(hipGraphLaunch(gexec, stream));
      } else if (timing) {
        int rc = launch_step(true);
        if (rc) return rc;
      } else {
        int rc = launch_step(false);
        if (rc) return rc;
      }
      ++steps_launched;
    }
    return PA_OK;
  }

  int create_stream() {
    int rc = prepare_kernels_once();
    if (rc) return rc;
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipEventCreate(&ev0));
    HIPCHECK(hipEventCreate(&ev1));
    return PA_OK;
  }
};

// Copy step records and state back; compute the results rows like SSC:620-630 and 679-690.
// Error bits of a run's state once its step records and state are on the host.
int check_fetched(RunHost& h) {
  if (h.st_h.err)
    return fail(PA_ERR_HIP, "device error bits " + std::to_string(h.st_h.err) +
                                " (1: a balanced-sweep hand-over wait timed out; 2: a sweep launch "
                                "shape left tasks unswept)");
  if (h.st_h.overflow)
    return fail(PA_ERR_CAPACITY, "population outgrew its buffers");
  h.results_valid = true;
  return PA_OK;
}

int fetch(RunHost& h, hipStream_t s) {
  h.out_h.resize(h.n_steps + 1);
  HIPCHECK(hipMemcpyAsync(h.out_h.data(), h.desc.out, (h.n_steps + 1) * sizeof(StepOut),
                          hipMemcpyDeviceToHost, s));
  HIPCHECK(hipMemcpyAsync(&h.st_h, h.desc.st, sizeof(RunState), hipMemcpyDeviceToHost, s));
  HIPCHECK(hipStreamSynchronize(s));
  return check_fetched(h);
}

int rows_of(const RunHost& h, pa_row* rows, int max_rows, int* nrows) {
  const int done = h.st_h.step - 1;  // completed steps
  const int N = h.N;
  double sumlnQ = 0.0;
  const int n = std::min(done + 1, max_rows);
  for (int i = 0; i < n; ++i) {
    pa_row& row = rows[i];
    memset(&row, 0, sizeof(row));
    row.B = h.B[i];
    if (i == 0) {
      // the reference prints never-assigned values for row 0 (they are 0 in its outputs)
      row.BF = (-N * std::log(2.0)) / N;
      row.R = h.p.Rinit;
      continue;
    }
    const StepOut& o = h.out_h[i];
    const double R = (double)o.R;
    const double E = o.sums[0] / o.R;
    const double C = (o.sums[1] / o.R - E * E) * h.B[i] * h.B[i];
    const double M = o.sums[2] / o.R;
    const double M2 = o.sums[3] / o.R;
    const double M4 = o.sums[4] / o.R;
    (void)R;
    sumlnQ -= o.lnQ;
    const double BF = -N * std::log(2.0) + sumlnQ;
    const double S = h.B[i] * E - BF;
    row.E = E / N;
    row.C = C / N;
    row.M = M / N;
    row.M2 = M2 / N / N;
    row.M4 = M4 / N / N / N / N;
    row.BF = BF / N;
    row.S = S / N;
    row.R = o.R;
    row.Rprev = o.Rprev;
    row.lnQ = o.lnQ;
    row.q = o.q;
    for (int k = 0; k < 5; ++k) row.sums[k] = o.sums[k];
    row.sumE = o.sumE;
    row.sumM = o.sumM;
    row.roff_hash = o.roff_hash;
    row.spin_hash = o.spin_hash;
    row.IS_after = o.IS_after;
  }
  *nrows = n;
  return PA_OK;
}

}  // namespace

// ============================== C ABI ==============================
struct pa_run {
  Batch batch;
  std::unique_ptr<RunHost> host;
  double cap_sigma = 10.0;
};

struct pa_batch {
  Batch batch;
};

extern "C" {

const char* pa_last_error(void) { return g_err.c_str(); }
int pa_version(void) { return 1; }

int pa_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes) {
  hipDeviceProp_t prop;
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  HIPCHECK(hipGetDeviceProperties(&prop, dev));
  if (name && name_len) {
    // some drivers leave the marketing name empty: describe the device by architecture,
    // compute units and engine clock instead
    if (prop.name[0])
      snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    else
      snprintf(name, name_len, "AMD %s, %d CUs, %d MHz", prop.gcnArchName,
               prop.multiProcessorCount, prop.clockRate / 1000);
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
  return PA_OK;
}

int pa_bandwidth_probe(size_t bytes, int iters, double* gbps) {
  if (bytes < 16 || iters < 1 || !gbps) return fail(PA_ERR_INVALID, "bad probe arguments");
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  HIPCHECK(hipMalloc(&a, bytes));
  HIPCHECK(hipMalloc(&b, bytes));
  HIPCHECK(hipMemset(a, 1, bytes));
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  HIPCHECK(pa_launch_copy(a, b, bytes, nullptr));
  HIPCHECK(hipEventRecord(e0, nullptr));
  for (int i = 0; i < iters; ++i) HIPCHECK(pa_launch_copy(a, b, bytes, nullptr));
  HIPCHECK(hipEventRecord(e1, nullptr));
  HIPCHECK(hipEventSynchronize(e1));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2.0 * bytes * iters / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return PA_OK;
}

int pa_schedule_steps(double Binit, double dB, double Bfin) {
  if (!(dB > 0.0)) return -1;
  return (int)schedule(Binit, dB, Bfin, 0).size() - 1;
}

size_t pa_run_footprint(const pa_params* p) {
  RunHost h;
  if (!p || plan_run(*p, h, 10.0)) return 0;
  return layout(h, nullptr);
}

static int run_create_sigma(const pa_params* p, double sigma, pa_run** out) {
  if (!p || !out) return fail(PA_ERR_INVALID, "null argument");
  auto r = std::make_unique<pa_run>();
  r->cap_sigma = sigma;
  r->host = std::make_unique<RunHost>();
  int rc = plan_run(*p, *r->host, sigma);
  if (rc) return rc;
  if ((rc = r->batch.create_stream())) return rc;
  if ((rc = materialize(*r->host))) return rc;
  r->batch.runs.push_back(r->host.get());
  if ((rc = r->batch.build_tables())) return rc;
  HIPCHECK(hipEventRecord(r->batch.ev0, r->batch.stream));
  if ((rc = r->batch.launch_init())) return rc;
  *out = r.release();
  return PA_OK;
}

int pa_run_create(const pa_params* p, pa_run** out) { return run_create_sigma(p, 10.0, out); }

int pa_run_step(pa_run* r, int nsteps) {
  if (!r) return fail(PA_ERR_INVALID, "null run");
  RunHost& h = *r->host;
  if (h.adaptive) {
    // one step at a time: each step's beta depends on the population after the last one
    h.results_valid = false;
    for (int todo = nsteps < 0 ? INT32_MAX : nsteps; todo > 0; --todo) {
      if (h.adaptive_done || r->batch.steps_launched >= h.n_steps) break;
      const int i = r->batch.steps_launched + 1;
      if (i >= 2) {
        int rc = r->batch.adaptive_next(h, i);
        if (rc) return rc;
        if (h.adaptive_done) break;
      }
      int rc = r->batch.run_steps(1);
      if (rc) return rc;
    }
    return PA_OK;
  }
  const int remaining = r->host->n_steps - r->batch.steps_launched;
  const int k = nsteps < 0 ? remaining : std::min(nsteps, remaining);
  r->host->results_valid = false;
  return r->batch.run_steps(std::max(k, 0));
}

int pa_run_sync(pa_run* r) {
  if (!r) return fail(PA_ERR_INVALID, "null run");
  HIPCHECK(hipEventRecord(r->batch.ev1, r->batch.stream));
  HIPCHECK(hipStreamSynchronize(r->batch.stream));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, r->batch.ev0, r->batch.ev1));
  r->batch.total_ms = ms;
  int rc = r->batch.collect_timing();
  if (rc) return rc;
  return fetch(*r->host, r->batch.stream);
}

int pa_run_steps_done(pa_run* r, int* done, int* total) {
  if (!r) return fail(PA_ERR_INVALID, "null run");
  int rc = pa_run_sync(r);
  if (rc) return rc;
  if (done) *done = r->host->st_h.step - 1;
  // adaptive: the number of steps is known once the schedule has passed Bfin
  if (total) *total = r->host->adaptive && r->host->adaptive_done ? (int)r->host->B.size() - 1
                                                                 : r->host->n_steps;
  return PA_OK;
}

int pa_run_results(pa_run* r, pa_row* rows, int max_rows, int* nrows) {
  if (!r || !rows || !nrows) return fail(PA_ERR_INVALID, "null argument");
  if (!r->host->results_valid) {
    int rc = pa_run_sync(r);
    if (rc) return rc;
  }
  return rows_of(*r->host, rows, max_rows, nrows);
}

int pa_run_final_sequence(pa_run* r, uint64_t* is) {
  if (!r || !is) return fail(PA_ERR_INVALID, "null argument");
  if (!r->host->results_valid) {
    int rc = pa_run_sync(r);
    if (rc) return rc;
  }
  *is = r->host->st_h.IS;
  return PA_OK;
}

int pa_run_population(pa_run* r, void* dst, size_t* bytes) {
  if (!r || !bytes) return fail(PA_ERR_INVALID, "null argument");
  int rc = pa_run_sync(r);
  if (rc) return rc;
  const RunHost& h = *r->host;
  const int R = h.st_h.R;
  const size_t unit = h.p.coding == PA_SSC ? (size_t)h.N : (size_t)h.N * (h.w / 8);
  const size_t units = h.p.coding == PA_SSC ? (size_t)R : (size_t)ceil_div(R, h.w);
  const size_t need = unit * units;
  if (!dst) { *bytes = need; return PA_OK; }
  if (*bytes < need) return fail(PA_ERR_INVALID, "buffer too small");
  HIPCHECK(hipMemcpy(dst, h.desc.pop[h.st_h.buf], need, hipMemcpyDeviceToHost));
  *bytes = need;
  return PA_OK;
}

int pa_run_array(pa_run* r, int which, void* dst, size_t* bytes) {
  if (!r || !bytes) return fail(PA_ERR_INVALID, "null argument");
  int rc = pa_run_sync(r);
  if (rc) return rc;
  const RunHost& h = *r->host;
  const int R = h.st_h.R;
  const int units = h.p.coding == PA_SSC ? R : ceil_div(R, h.w);
  const void* src = nullptr;
  size_t need = 0;
  const size_t members = h.p.coding == PA_SSC ? (size_t)R : (size_t)units * h.w;
  switch (which) {
    case 0: src = h.desc.IE; need = members * 4; break;
    case 1: src = h.desc.M; need = members * 4; break;
    case 2: case 3: {
      // arrays of the last resampling: indexed by the parents (previous population)
      const int Rp = h.st_h.step > 1 ? h.out_h[h.st_h.step - 1].Rprev : R;
      const size_t pm = h.p.coding == PA_SSC ? (size_t)Rp : (size_t)ceil_div(Rp, h.w) * h.w;
      src = which == 2 ? (const void*)h.desc.Roff : (const void*)h.desc.pre;
      need = pm * 4;
      break;
    }
    case 4: src = h.desc.amask[h.st_h.buf]; need = (size_t)units * 8; break;
    default: return fail(PA_ERR_INVALID, "which");
  }
  if (!dst) { *bytes = need; return PA_OK; }
  if (*bytes < need) return fail(PA_ERR_INVALID, "buffer too small");
  HIPCHECK(hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
  *bytes = need;
  return PA_OK;
}

int pa_run_timing(pa_run* r, double* sweep_ms, double* total_ms) {
  if (!r) return fail(PA_ERR_INVALID, "null run");
  if (sweep_ms) *sweep_ms = r->batch.sweep_ms;
  if (total_ms) *total_ms = r->batch.total_ms;
  return PA_OK;
}

int pa_run_kernel_times(pa_run* r, double* resample_ms, double* sweep_ms, int* steps) {
  if (!r) return fail(PA_ERR_INVALID, "null run");
  if (resample_ms) *resample_ms = r->batch.resample_ms;
  if (sweep_ms) *sweep_ms = r->batch.sweep_ms;
  if (steps) *steps = r->batch.timed_steps;
  return PA_OK;
}

int pa_run_set_timing(pa_run* r, int on) {
  if (!r) return fail(PA_ERR_INVALID, "null run");
  r->batch.timing = on != 0;
  return PA_OK;
}

int pa_run_set_graph(pa_run* r, int on) {
  if (!r) return fail(PA_ERR_INVALID, "null run");
  r->batch.use_graph = on != 0;
  return PA_OK;
}

int pa_run_destroy(pa_run* r) {
  delete r;
  return PA_OK;
}

int pa_run_histogram(pa_run* r, int32_t* hist, int nbins) {
  if (!r || !hist) return fail(PA_ERR_INVALID, "null argument");
  const RunHost& h = *r->host;
  if (!h.desc.hist) return fail(PA_ERR_STATE, "run created without mhr = 1");
  if (nbins < h.N + 1) return fail(PA_ERR_INVALID, "histogram needs N + 1 bins");
  HIPCHECK(hipStreamSynchronize(r->batch.stream));
  HIPCHECK(hipMemcpy(hist, h.desc.hist, (size_t)(h.N + 1) * 4, hipMemcpyDeviceToHost));
  return PA_OK;
}

// MHR analysis, PAisingSSC.cu:705-764 (identical in PAisingMSC.cu:855-914).  Host FP64 with
// the C library's exp, log and pow, in the reference's order of operations.
int pa_mhr_analysis(int N, int nB, const double* B, double* BF, const int32_t* R,
                    const int32_t* hist, int niter, double* out) {
  if (N < 1 || nB < 1 || !B || !BF || !R || !hist || niter < 1 || !out)
    return fail(PA_ERR_INVALID, "bad MHR arguments");
  const int nk = N + 1;
  std::vector<double> Ei(nk), lnOmega(nk), Sigma(nB), expOm(nk), Om(nk);
  std::vector<char> rel(nk);
  for (int k = 0; k < nk; ++k) Ei[k] = 4 * k - 2 * N;  // SSC:461-463
  const int cols = 1 + 3 * niter;
  for (int i = 0; i < nB; ++i) out[(size_t)i * cols] = B[i];
  for (int l = 0; l < niter; ++l) {
    // density of states from the summed histogram (SSC:707-731)
    for (int k = 0; k < nk; ++k) {
      double mS = BF[0] - B[0] * Ei[k];
      Sigma[0] = mS;
      for (int i = 1; i < nB; ++i) {
        Sigma[i] = BF[i] - B[i] * Ei[k];
        if (mS < Sigma[i]) mS = Sigma[i];
      }
      double sD = 0;
      for (int i = 0; i < nB; ++i) sD += R[i] * std::exp(Sigma[i] - mS);
      if (hist[k] == 0 || sD == 0) {
        rel[k] = 0;
        lnOmega[k] = 0;
      } else {
        rel[k] = 1;
        lnOmega[k] = std::log((double)hist[k]) - mS - std::log(sD);
      }
    }
    // reweighted observables per temperature (SSC:733-764); the maximum runs over every
    // bin, relevant or not, as in the reference
    for (int i = 0; i < nB; ++i) {
      double mOm = lnOmega[0] - B[i] * Ei[0];
      for (int k = 0; k < nk; ++k) {
        Om[k] = lnOmega[k] - B[i] * Ei[k];
        if (mOm < Om[k]) mOm = Om[k];
      }
      double p = 0;
      for (int k = 0; k < nk; ++k) {
        expOm[k] = std::exp(Om[k] - mOm);
        if (rel[k]) p += expOm[k];
      }
      double s = 0;
      for (int k = 0; k < nk; ++k)
        if (rel[k]) s += Ei[k] * expOm[k];
      const double E = s / p / N;
      const double F = -mOm - std::log(p);
      BF[i] = F;
      s = 0;
      for (int k = 0; k < nk; ++k)
        if (rel[k]) s += std::pow(Ei[k] - E * N, 2) * expOm[k];
      const double C = B[i] * B[i] * s / p / N;
      double* o = out + (size_t)i * cols + 1 + 3 * l;
      o[0] = E;
      o[1] = C;
      o[2] = F / N;
    }
  }
  return PA_OK;
}

int pa_run_all(const pa_params* p, pa_row* rows, int max_rows, int* nrows, uint64_t* final_is) {
  return pa_run_all_hist(p, rows, max_rows, nrows, final_is, nullptr, 0);
}

int pa_run_all_hist(const pa_params* p, pa_row* rows, int max_rows, int* nrows,
                    uint64_t* final_is, int32_t* hist, int nbins) {
  double sigma = 10.0;
  for (int attempt = 0; attempt < 4; ++attempt, sigma *= 4.0) {
    pa_run* r = nullptr;
    int rc = run_create_sigma(p, sigma, &r);
    if (rc) return rc;
    rc = pa_run_step(r, -1);
    if (!rc) rc = pa_run_sync(r);
    if (rc == PA_ERR_CAPACITY) { pa_run_destroy(r); continue; }
    if (!rc) rc = rows_of(*r->host, rows, max_rows, nrows);
    if (!rc && final_is) *final_is = r->host->st_h.IS;
    if (!rc && hist) rc = pa_run_histogram(r, hist, nbins);
    pa_run_destroy(r);
    return rc;
  }
  return fail(PA_ERR_CAPACITY, "population outgrew its buffers repeatedly");
}

// ---------------- autotuner ----------------
// Times each candidate sweep variant (slice width, workgroup size, balanced or not) on a
// short trial of the real run (2 steps, the second timed with HIP events around the sweep
// launches) and caches the fastest for runs of the same shape.  All variants produce the
// same bits, so tuning only changes speed.
int pa_autotune(const pa_params* p, pa_params* tuned, double* best_sweep_ms) {
  if (!p || !tuned) return fail(PA_ERR_INVALID, "null argument");
  RunHost probe;
  int rc = plan_run(*p, probe, 10.0);
  if (rc) return rc;
  std::vector<std::array<int, 3>> cands;
  const int nts[4] = {128, 256, 512, 1024};
  if (p->coding == PA_SSC) {
    // the plan's own shape, then workgroup sizes of the packed-byte kernels (where the
    // lattice takes them) and of the site-per-thread kernels (slice_bits = 1); shapes a
    // kernel does not take fail at creation and are skipped
    cands.push_back({0, 0, 0});
    for (int nt : {64, 128, 256, 512, 1024}) cands.push_back({0, nt, 0});
    const int Nh = p->L * p->L / 2;
    if (p->L % 8 == 0 && !(Nh % 512 == 0 && p->L >= 64 && p->L <= 256)) {
      // the packed kernel where the plan prefers the site kernels (small populations)
      cands.push_back({2, 0, 0});
      for (int nt : {64, 128, 256, 512}) cands.push_back({2, nt, 0});
    }
    cands.push_back({1, 0, 0});
    for (int nt : {128, 256, 512}) cands.push_back({1, nt, 0});
  }
  for (int nt : nts) {
    if (p->coding == PA_SSC) break;
    for (int sw = p->msbits; sw >= 8; sw /= 2) {
      // sw = 32 may still fit with one sublattice in LDS (the plan decides)
      if ((size_t)p->L * p->L * (sw / 8) + 2048 > kLdsLimit && sw != 32) continue;
      cands.push_back({sw, nt, -1});
      if (nt <= 512 && p->msbits >= 32) cands.push_back({sw, nt, 1});
    }
  }
  double best = 1e300;
  pa_params bestp = *p;
  for (const auto& c : cands) {
    pa_params q = *p;
    q.slice_bits = c[0];
    q.threads = c[1];
    q.balance = c[2];
    q.max_steps = 2;
    q.trace = 0;
    pa_run* r = nullptr;
    rc = pa_run_create(&q, &r);
    if (rc) continue;  // variant not valid for this shape
    double ms = 1e300;
    if (pa_run_step(r, 1) == PA_OK && pa_run_sync(r) == PA_OK) {
      r->batch.timing = true;
      r->batch.sweep_ms = 0;
      if (pa_run_step(r, 1) == PA_OK && pa_run_sync(r) == PA_OK) ms = r->batch.sweep_ms;
    }
    pa_run_destroy(r);
    if (ms < best) {
      best = ms;
      bestp = q;
    }
  }
  if (best >= 1e299) return fail(PA_ERR_INVALID, "no sweep variant ran for this shape");
  *tuned = *p;
  tuned->slice_bits = bestp.slice_bits;
  tuned->threads = bestp.threads;
  tuned->balance = bestp.balance;
  tune_cache()[tune_key(*p)] = TuneVal{bestp.slice_bits, bestp.threads, bestp.balance, best};
  if (best_sweep_ms) *best_sweep_ms = best;
  return PA_OK;
}

int pa_autotune_clear(void) {
  tune_cache().clear();
  return PA_OK;
}

// ---------------- single stages on caller data ----------------
// Each stage entry point builds a one-step scratch run, fills its state and arrays from the
// caller's host arrays, launches the kernels of that part of a step (the same launches as a
// step of a run) and copies the results back.
namespace {

struct StageRun {
  RunHost h;
  Batch b;  // declared after h: destroyed first
};

// units (replicas or groups) and members (replicas or group lanes) of a population of R
int stage_units(const RunHost& h, int R) { return h.p.coding == PA_SSC ? R : ceil_div(R, h.w); }
size_t stage_members(const RunHost& h, int R) {
  return h.p.coding == PA_SSC ? (size_t)R : (size_t)ceil_div(R, h.w) * h.w;
}
size_t stage_unit_bytes(const RunHost& h) {
  return h.p.coding == PA_SSC ? (size_t)h.N : (size_t)h.N * (h.w / 8);
}

// Scratch run with room for max(R, cap) members, step 1 at inverse temperature beta with
// increment dB; the caller's Rinit is kept for the copy-count ratio.
int stage_open(const pa_params* p, int R, int cap, double beta, double dB, StageRun& s) {
  if (!p || R < 1 || cap < 1) return fail(PA_ERR_INVALID, "bad stage arguments");
  pa_params q = *p;
  q.Rinit = std::max(R, cap);
  q.Binit = 0.0;
  q.dB = 1.0;
  q.Bfin = 1.0;
  q.max_steps = 1;
  q.adaptive = 0;
  q.schedule = nullptr;
  q.schedule_len = 0;
  q.mhr = 0;
  q.trace = 0;
  int rc = plan_run(q, s.h, 10.0);
  if (rc) return rc;
  if ((rc = materialize(s.h))) return rc;
  s.h.desc.Rinit = p->Rinit;
  const double bt[2] = {0.0, beta}, db[2] = {0.0, dB};
  uint64_t thr[3];
  step_thresholds(s.h, beta, thr);
  HIPCHECK(hipMemcpy((void*)s.h.desc.beta, bt, sizeof bt, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy((void*)s.h.desc.dbeta, db, sizeof db, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy((void*)(s.h.desc.thr + 3), thr, sizeof thr, hipMemcpyHostToDevice));
  if ((rc = s.b.create_stream())) return rc;
  s.b.runs = {&s.h};
  return s.b.build_tables();
}

int stage_state(StageRun& s, int R, int Rnew, int buf, uint64_t IS, double Emean) {
  RunState st{};
  st.step = 1;
  st.R = R;
  st.Rnew = Rnew;
  st.buf = buf;
  st.IS = IS;
  st.Emean = Emean;
  st.active = 1;
  HIPCHECK(hipMemcpy(s.h.desc.st, &st, sizeof st, hipMemcpyHostToDevice));
  return PA_OK;
}

int stage_sync(StageRun& s, RunState* st) {
  HIPCHECK(hipStreamSynchronize(s.b.stream));
  HIPCHECK(hipMemcpy(st, s.h.desc.st, sizeof *st, hipMemcpyDeviceToHost));
  if (st->err) return fail(PA_ERR_HIP, "device error bits " + std::to_string(st->err));
  if (st->overflow) return fail(PA_ERR_CAPACITY, "children exceed the stage's capacity");
  return PA_OK;
}

// one byte per lane <-> 64-bit lane mask per group
int stage_put_masks(const RunHost& h, int which, const uint8_t* active, int R) {
  const int rg = ceil_div(R, h.w);
  std::vector<uint64_t> m(rg, 0);
  for (int g = 0; g < rg; ++g)
    for (int k = 0; k < h.w; ++k)
      if (active ? active[(size_t)g * h.w + k] != 0 : (int64_t)g * h.w + k < R) m[g] |= 1ull << k;
  HIPCHECK(hipMemcpy(h.desc.amask[which], m.data(), m.size() * 8, hipMemcpyHostToDevice));
  return PA_OK;
}
int stage_get_masks(const RunHost& h, int which, uint8_t* active, int R) {
  const int rg = ceil_div(R, h.w);
  std::vector<uint64_t> m(rg, 0);
  HIPCHECK(hipMemcpy(m.data(), h.desc.amask[which], m.size() * 8, hipMemcpyDeviceToHost));
  for (int g = 0; g < rg; ++g)
    for (int k = 0; k < h.w; ++k) active[(size_t)g * h.w + k] = (uint8_t)((m[g] >> k) & 1);
  return PA_OK;
}

int stage_sweep(const pa_params* p, int R, uint64_t is, double beta, void* pop, int32_t* IE,
                int32_t* M) {
  if (!pop) return fail(PA_ERR_INVALID, "null population");
  StageRun s;
  int rc = stage_open(p, R, R, beta, 0.0, s);
  if (rc) return rc;
  const RunHost& h = s.h;
  const size_t bytes = stage_unit_bytes(h) * stage_units(h, R);
  HIPCHECK(hipMemcpy(h.desc.pop[0], pop, bytes, hipMemcpyHostToDevice));
  // the sweep reads the population the resampling wrote (buffer buf ^ 1) and draws from
  // subsequence IS + (parents' tau draws) + t + 128 unit; with no parents that is `is`
  if ((rc = stage_state(s, 0, R, 1, is, 0.0))) return rc;
  if ((rc = s.b.launch_sweeps())) return rc;
  RunState st;
  if ((rc = stage_sync(s, &st))) return rc;
  HIPCHECK(hipMemcpy(pop, h.desc.pop[0], bytes, hipMemcpyDeviceToHost));
  const size_t mem = stage_members(h, R);
  if (IE) HIPCHECK(hipMemcpy(IE, h.desc.IE, mem * 4, hipMemcpyDeviceToHost));
  if (M) HIPCHECK(hipMemcpy(M, h.desc.M, mem * 4, hipMemcpyDeviceToHost));
  return PA_OK;
}

}  // namespace

extern "C" {

int pa_stage_sweep(const pa_params* p, int R, uint64_t is, double beta, void* pop, int32_t* IE,
                   int32_t* M) {
  return stage_sweep(p, R, is, beta, pop, IE, M);
}

int pa_stage_energy(const pa_params* p, int R, const void* pop, int32_t* IE, int32_t* M) {
  if (!p || !pop) return fail(PA_ERR_INVALID, "null argument");
  pa_params q = *p;
  q.eqsweeps = 0;
  // zero sweeps leave the population unchanged; work on a copy to keep `pop` const
  const size_t unit = p->coding == PA_SSC ? (size_t)p->L * p->L : (size_t)p->L * p->L * (p->msbits / 8);
  const size_t units = p->coding == PA_SSC ? (size_t)R : (size_t)ceil_div(R, std::max(1, p->msbits));
  std::vector<uint8_t> tmp((const uint8_t*)pop, (const uint8_t*)pop + unit * units);
  return stage_sweep(&q, R, 0, 0.0, tmp.data(), IE, M);
}

int pa_stage_copy_counts(const pa_params* p, int R, const int32_t* IE, const uint8_t* active,
                         double dB, double Emean, uint64_t is, double* q, double* lnQ,
                         uint32_t* Roff, uint32_t* prefix, int* Rnew) {
  if (!p || !IE) return fail(PA_ERR_INVALID, "null argument");
  StageRun s;
  int rc = stage_open(p, R, p->Rinit, 0.0, dB, s);
  if (rc) return rc;
  const RunHost& h = s.h;
  const size_t mem = stage_members(h, R);
  HIPCHECK(hipMemcpy(h.desc.IE, IE, mem * 4, hipMemcpyHostToDevice));
  if (h.p.coding == PA_MSC && (rc = stage_put_masks(h, 0, active, R))) return rc;
  if ((rc = stage_state(s, R, 0, 0, is, Emean))) return rc;
  if ((rc = s.b.launch_counts()) || (rc = s.b.launch_scan())) return rc;
  RunState st;
  if ((rc = stage_sync(s, &st))) return rc;
  StepOut o;
  HIPCHECK(hipMemcpy(&o, h.desc.out + 1, sizeof o, hipMemcpyDeviceToHost));
  if (q) *q = o.q;
  if (lnQ) *lnQ = o.lnQ;
  if (Roff) HIPCHECK(hipMemcpy(Roff, h.desc.Roff, mem * 4, hipMemcpyDeviceToHost));
  if (prefix) HIPCHECK(hipMemcpy(prefix, h.desc.pre, mem * 4, hipMemcpyDeviceToHost));
  if (Rnew) *Rnew = st.Rnew;
  return PA_OK;
}

int pa_stage_resample(const pa_params* p, int R, const void* pop, const uint8_t* active,
                      const uint32_t* Roff, void* pop_new, uint8_t* active_new, int* Rnew) {
  if (!p || !pop || !Roff || !pop_new) return fail(PA_ERR_INVALID, "null argument");
  const int w = p->coding == PA_SSC ? 1 : p->msbits;
  if (w < 1) return fail(PA_ERR_INVALID, "msbits");
  const size_t mem = p->coding == PA_SSC ? (size_t)R : (size_t)ceil_div(R, w) * w;
  // the copy-count stage leaves per-1024 chunk totals for the scan; recreate them
  std::vector<uint32_t> cblk(ceil_div((int)mem, kRedBlock), 0);
  int64_t total = 0;
  for (size_t n = 0; n < mem; ++n) {
    const bool act = p->coding == PA_SSC || !active || active[n] != 0;
    const uint32_t r = act ? Roff[n] : 0u;
    cblk[n / kRedBlock] += r;
    total += r;
  }
  if (total < 1 || total > INT32_MAX / 2) return fail(PA_ERR_INVALID, "copy counts sum");
  StageRun s;
  int rc = stage_open(p, R, (int)total, 0.0, 0.0, s);
  if (rc) return rc;
  const RunHost& h = s.h;
  HIPCHECK(hipMemcpy(h.desc.pop[0], pop, stage_unit_bytes(h) * stage_units(h, R),
                     hipMemcpyHostToDevice));
  std::vector<uint32_t> ro(Roff, Roff + mem);
  if (p->coding == PA_MSC && active)
    for (size_t n = 0; n < mem; ++n)
      if (!active[n]) ro[n] = 0;
  HIPCHECK(hipMemcpy(h.desc.Roff, ro.data(), mem * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(h.desc.cblk, cblk.data(), cblk.size() * 4, hipMemcpyHostToDevice));
  if ((rc = stage_state(s, R, 0, 0, 0, 0.0))) return rc;
  if ((rc = s.b.launch_scan()) || (rc = s.b.launch_resample())) return rc;
  RunState st;
  if ((rc = stage_sync(s, &st))) return rc;
  HIPCHECK(hipMemcpy(pop_new, h.desc.pop[1], stage_unit_bytes(h) * stage_units(h, st.Rnew),
                     hipMemcpyDeviceToHost));
  if (h.p.coding == PA_MSC && active_new && (rc = stage_get_masks(h, 1, active_new, st.Rnew)))
    return rc;
  if (Rnew) *Rnew = st.Rnew;
  return PA_OK;
}

int pa_stage_averages(const pa_params* p, int R, const int32_t* IE, const int32_t* M,
                      const uint8_t* active, double sums[5], int64_t* sumE, int64_t* sumM) {
  if (!p || !IE || !M || !sums) return fail(PA_ERR_INVALID, "null argument");
  StageRun s;
  int rc = stage_open(p, R, R, 0.0, 0.0, s);
  if (rc) return rc;
  const RunHost& h = s.h;
  const size_t mem = stage_members(h, R);
  HIPCHECK(hipMemcpy(h.desc.IE, IE, mem * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(h.desc.M, M, mem * 4, hipMemcpyHostToDevice));
  // the averages read the new population: buffer buf ^ 1
  if (h.p.coding == PA_MSC && (rc = stage_put_masks(h, 1, active, R))) return rc;
  if ((rc = stage_state(s, R, R, 0, 0, 0.0))) return rc;
  if ((rc = s.b.launch_finish())) return rc;
  RunState st;
  if ((rc = stage_sync(s, &st))) return rc;
  StepOut o;
  HIPCHECK(hipMemcpy(&o, h.desc.out + 1, sizeof o, hipMemcpyDeviceToHost));
  for (int k = 0; k < 5; ++k) sums[k] = o.sums[k];
  if (sumE) *sumE = o.sumE;
  if (sumM) *sumM = o.sumM;
  return PA_OK;
}

}  // extern "C"

// ---------------- allocation-free sweep stage ----------------
// pa_stage_sweep_noalloc: the sweeps and energies of pa_stage_sweep on caller-owned device
// memory, for callers that cannot allocate (code under HIP graph capture, serving loops).
// The only device memory beyond the caller's population and energy arrays is the run's
// descriptor, launch table, state and thresholds: a fixed NoAllocTables record in the
// file-scope `scratch` array below (one copy per device, held until the process ends), written
// by a setup kernel from its by-value arguments -- no copy or memset call -- and read by the
// sweep launch that follows it on the same stream.  Calls that overlap on one device race on
// that record.
namespace {

struct NoAllocTables {
  RunDesc desc;
  int32_t offs[kKindCount * 2];  // launch table of one run (the per-unit and packed-sweep rows)
  RunState st;
  uint64_t thr[6];               // thresholds of step 1 at [3..5]
  double beta[2];
  double dbeta[2];
};
constexpr size_t kNoAllocScratchBytes = 4096;
static_assert(sizeof(NoAllocTables) <= kNoAllocScratchBytes, "scratch record exceeds the array");
static_assert(sizeof(NoAllocTables) % 4 == 0, "setup kernel copies 32-bit words");

static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

// The array's device address per device (the host symbol only names it), taken once per
// device with that device current.
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(void** out) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(PA_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(PA_ERR_HIP, "no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return PA_OK;
}

__global__ __launch_bounds__(64) void k_noalloc_setup(NoAllocTables t, NoAllocTables* dst) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&t);
  uint32_t* d = reinterpret_cast<uint32_t*>(dst);
  for (int i = threadIdx.x; i < (int)(sizeof(NoAllocTables) / 4); i += blockDim.x) d[i] = src[i];
}

}  // namespace

extern "C" int pa_stage_sweep_noalloc(const pa_params* p, int R, uint64_t is, double beta,
                                      void* pop, int32_t* IE, int32_t* M, void* stream) {
  if (!p || !pop || !IE || !M) return fail(PA_ERR_INVALID, "null argument");
  if (R < 1) return fail(PA_ERR_INVALID, "R must be at least 1");
  if (p->adaptive || p->mhr || p->schedule_len)
    return fail(PA_ERR_INVALID, "the sweep stage takes no schedule, adaptive or MHR options");
  pa_params q = *p;
  q.Rinit = R;
  q.Binit = 0.0;
  q.dB = 1.0;
  q.Bfin = 1.0;
  q.max_steps = 1;
  q.trace = 0;
  q.schedule = nullptr;
  // makespan-balanced sweeps hand units over through per-unit flags that would need
  // scratch proportional to R; every sweep variant gives the same bits
  q.balance = -1;
  RunHost h;
  int rc = plan_run(q, h, 10.0);
  if (rc) return rc;
  if (sizeof(NoAllocTables) > kNoAllocScratchBytes)
    return fail(PA_ERR_INVALID, "scratch record exceeds the " + std::to_string(kNoAllocScratchBytes) +
                                    "-byte array");
  void* ptr = nullptr;
  if ((rc = noalloc_scratch(&ptr))) return rc;
  if ((rc = prepare_kernels_once())) return rc;
  NoAllocTables* dev = reinterpret_cast<NoAllocTables*>(ptr);
  const int units = h.p.coding == PA_SSC ? R : ceil_div(R, h.w);
  h.cap_units = units;
  desc_scalars(h);
  NoAllocTables t;
  memset(&t, 0, sizeof t);
  t.desc = h.desc;
  t.desc.pop[0] = t.desc.pop[1] = pop;
  t.desc.IE = IE;
  t.desc.M = M;
  t.desc.st = &dev->st;
  t.desc.thr = dev->thr;
  t.desc.beta = dev->beta;
  t.desc.dbeta = dev->dbeta;
  t.offs[kKUnits * 2 + 0] = 0;
  t.offs[kKUnits * 2 + 1] = units;
  t.offs[kKSweep * 2 + 0] = 0;
  t.offs[kKSweep * 2 + 1] = h.blocks(kKSweep);
  // step 1 of a run whose sweep streams start at `is` (no parents: R = 0), reading the
  // population from buffer buf ^ 1 = 0, as pa_stage_sweep sets it up
  t.st.step = 1;
  t.st.R = 0;
  t.st.Rnew = R;
  t.st.buf = 1;
  t.st.IS = is;
  t.st.active = 1;
  step_thresholds(h, beta, &t.thr[3]);
  t.beta[1] = beta;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(64), 0, s, t, dev);
  HIPCHECK(hipGetLastError());
  const size_t lds = h.p.coding == PA_SSC ? (size_t)h.N
                     : h.msc_unit_lds();
  if (h.p.coding == PA_MSC && h.half)
    HIPCHECK(pa_launch_step(16, PA_MSC, h.w, 32, 1024, units * (h.w / 32), lds, s, &dev->desc,
                            dev->offs, 1));
  else if (h.p.coding == PA_SSC && (h.path == 8 || h.path == 9))
    HIPCHECK(pa_launch_step(4, PA_SSC, 1, h.path, h.nt, h.blocks(kKSweep), h.sweep_lds(), s,
                            &dev->desc, dev->offs, 1));
  else
    HIPCHECK(pa_launch_step(4, h.p.coding, h.w, h.p.coding == PA_SSC ? h.path : h.sw, h.nt,
                            units * (h.p.coding == PA_SSC ? 1 : h.w / h.sw), lds, s, &dev->desc,
                            dev->offs, 1));
  return PA_OK;
}

// ---------------- batch ----------------
int pa_batch_create(const pa_params* p, int nruns, int use_graph, pa_batch** out) {
  if (!p || nruns < 1 || !out) return fail(PA_ERR_INVALID, "bad batch arguments");
  auto b = std::make_unique<pa_batch>();
  int rc;
  if ((rc = b->batch.create_stream())) return rc;
  // MSC runs of one word width share the sweep launches; single-spin runs count the other
  // single-spin runs' replicas x sites
  auto msc_work = [&](int i) {
    return p[i].coding == PA_MSC && p[i].msbits > 0
               ? (int64_t)ceil_div(p[i].Rinit, p[i].msbits) * p[i].L * p[i].L : (int64_t)0;
  };
  for (int i = 0; i < nruns; ++i) {
    if (p[i].adaptive) return fail(PA_ERR_INVALID, "adaptive runs take the plain path (pa_run_*)");
    int64_t co = 0;
    for (int j = 0; j < nruns; ++j) {
      if (j == i) continue;
      if (p[i].coding == PA_SSC)
        co += p[j].coding == PA_SSC ? (int64_t)p[j].Rinit * p[j].L * p[j].L : 0;
      else if (p[j].msbits == p[i].msbits)
        co += msc_work(j);
    }
    auto h = std::make_unique<RunHost>();
    if ((rc = plan_run(p[i], *h, 10.0, co))) return rc;
    if ((rc = materialize(*h))) return rc;
    b->batch.runs.push_back(h.get());
    b->batch.owned.push_back(std::move(h));
  }
  b->batch.use_graph = use_graph != 0;
  if ((rc = b->batch.build_tables())) return rc;
  *out = b.release();
  return PA_OK;
}

int pa_batch_run(pa_batch* b) {
  if (!b) return fail(PA_ERR_INVALID, "null batch");
  Batch& B = b->batch;
  int rc;
  HIPCHECK(hipEventRecord(B.ev0, B.stream));
  if ((rc = B.launch_init())) return rc;
  if ((rc = B.run_steps(B.max_steps))) return rc;
  HIPCHECK(hipEventRecord(B.ev1, B.stream));
  HIPCHECK(hipStreamSynchronize(B.stream));
  float ms = 0;
  HIPCHECK(hipEventElapsedTime(&ms, B.ev0, B.ev1));
  B.total_ms = ms;
  for (RunHost* h : B.runs)
    if ((rc = fetch(*h, B.stream))) return rc;
  return PA_OK;
}

int pa_batch_results(pa_batch* b, int run, pa_row* rows, int max_rows, int* nrows) {
  if (!b || run < 0 || run >= (int)b->batch.runs.size()) return fail(PA_ERR_INVALID, "run index");
  if (!b->batch.runs[run]->results_valid) return fail(PA_ERR_STATE, "batch not run");
  return rows_of(*b->batch.runs[run], rows, max_rows, nrows);
}

int pa_batch_time(pa_batch* b, double* total_ms) {
  if (!b) return fail(PA_ERR_INVALID, "null batch");
  if (total_ms) *total_ms = b->batch.total_ms;
  return PA_OK;
}

int pa_batch_destroy(pa_batch* b) {
  delete b;
  return PA_OK;
}

}  // extern "C"

// ---------------- server: continuous batching under an HBM budget ----------------
// Requests wait in a FIFO queue.  Each tick picks up the results of runs whose last step
// has completed on the device, admits queued runs while their footprint fits the budget (and
// the resident-run limit), and advances the scheduled resident runs with one batched launch
// sequence.  Runs that join later start at their own step 1: every kernel reads its run's
// own state, so a run's results do not depend on which other runs share the device.
//
// Nothing on the tick path allocates device memory or drains the device: arenas and tables
// come from the server's pools (ServerMem), uploads are stream-ordered copies from pinned
// memory, and a run whose last step has been enqueued has its step records copied into
// pinned memory behind that step; a later tick finds the copy complete (an event query) and
// only then returns the run's arena to the pool.
namespace {

struct Request {
  int64_t ticket = 0;
  pa_params p{};
  size_t footprint = 0;
  std::unique_ptr<RunHost> h;
  int state = 0;  // 0 queued, 1 resident, 2 done, 3 failed, 4 last step enqueued (draining)
  int steps_left = 0;
  double credit = 0;  // serving: step credit carried between ticks
  int rc = PA_OK;
  std::string err;
  std::vector<pa_row> rows;
  double t_submit = 0, t_admit = 0, t_done = 0;
  void* pin = nullptr;  // draining: step records + state, copied behind fence `seq`
  int64_t seq = 0;
};

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct pa_server {
  size_t budget = 0;
  int max_runs = 0;
  size_t used = 0;   // footprints of the admitted runs whose arenas are not back in the pool
  ServerMem mem;     // before `batch`: the batch hands its tables back to these pools
  Batch batch;
  std::deque<int64_t> queue;
  std::map<int64_t, Request> reqs;
  std::vector<int64_t> resident;
  std::vector<int64_t> draining;
  int64_t next_ticket = 1;
  int completed = 0;
  int chunk = 32;
  int policy = PA_SERVE_SHORTEST;
  bool dirty = false;
  std::vector<int32_t> pause, quota;
  int64_t tick_seq[2] = {0, 0};  // fences of the last two ticks that enqueued work
  int64_t ticks = 0;
};

extern "C" {

int pa_server_create(size_t hbm_budget_bytes, int max_runs, pa_server** out) {
  if (!out || hbm_budget_bytes == 0 || max_runs < 1) return fail(PA_ERR_INVALID, "bad server arguments");
  auto s = std::make_unique<pa_server>();
  s->budget = hbm_budget_bytes;
  s->max_runs = max_runs;
  int rc = s->batch.create_stream();
  if (rc) return rc;
  s->mem.stream = s->batch.stream;
  // arenas only; the per-run footprints sum to at most the budget, fragmentation aside
  s->mem.dev_cap = hbm_budget_bytes + (size_t)256 * 4096;
  s->batch.mem = &s->mem;
  s->batch.use_graph = true;
  *out = s.release();
  return PA_OK;
}

int pa_server_reserve(pa_server* s, size_t bytes) {
  if (!s) return fail(PA_ERR_INVALID, "null server");
  bytes = std::min(bytes, s->mem.dev_cap);
  if (s->mem.dev.reserved() >= bytes) return PA_OK;
  HIPCHECK(s->mem.dev.grow(bytes - s->mem.dev.reserved()));
  return PA_OK;
}

int pa_server_memory(pa_server* s, size_t* reserved, int* device_allocations) {
  if (!s) return fail(PA_ERR_INVALID, "null server");
  if (reserved) *reserved = s->mem.dev.reserved() + s->mem.tab.reserved();
  if (device_allocations) *device_allocations = s->mem.dev.grows() + s->mem.tab.grows();
  return PA_OK;
}

int pa_server_set_chunk(pa_server* s, int steps_per_tick, int use_graph) {
  if (!s || steps_per_tick < 1) return fail(PA_ERR_INVALID, "bad chunk");
  s->chunk = steps_per_tick;
  s->batch.use_graph = use_graph != 0;
  return PA_OK;
}

int pa_server_set_policy(pa_server* s, int policy) {
  if (!s || (policy != PA_SERVE_SHARE && policy != PA_SERVE_SHORTEST))
    return fail(PA_ERR_INVALID, "bad policy");
  s->policy = policy;
  return PA_OK;
}

int pa_server_submit(pa_server* s, const pa_params* p, int64_t* ticket) {
  if (!s || !p || !ticket) return fail(PA_ERR_INVALID, "null argument");
  if (p->adaptive) return fail(PA_ERR_INVALID, "adaptive runs take the plain path (pa_run_*)");
  Request rq;
  rq.ticket = s->next_ticket++;
  rq.p = *p;
  rq.h = std::make_unique<RunHost>();
  // served runs share the chip with the other requests: widest slices (least total work)
  int rc = plan_run(*p, *rq.h, 10.0, (int64_t)1 << 40);
  if (rc) return rc;
  rq.footprint = layout(*rq.h, nullptr);
  if (rq.footprint > s->budget) return fail(PA_ERR_MEMORY, "request exceeds the HBM budget");
  rq.t_submit = now_s();
  *ticket = rq.ticket;
  s->queue.push_back(rq.ticket);
  s->reqs.emplace(rq.ticket, std::move(rq));
  return PA_OK;
}

}  // extern "C"

namespace {

// Results of the draining runs whose copies have completed (no wait).
void server_collect(pa_server* s) {
  const int64_t done = s->mem.poll();
  if (s->draining.empty()) return;
  const double tnow = now_s();
  std::vector<int64_t> keep;
  for (int64_t t : s->draining) {
    Request& rq = s->reqs[t];
    if (rq.seq > done) { keep.push_back(t); continue; }
    RunHost& h = *rq.h;
    h.out_h.resize(h.n_steps + 1);
    memcpy(h.out_h.data(), rq.pin, (h.n_steps + 1) * sizeof(StepOut));
    memcpy(&h.st_h, (const uint8_t*)rq.pin + (h.n_steps + 1) * sizeof(StepOut), sizeof(RunState));
    s->mem.pin.give(rq.pin);
    rq.pin = nullptr;
    rq.rc = check_fetched(h);
    if (rq.rc == PA_OK) {
      rq.rows.resize(h.n_steps + 1);
      int n = 0;
      rows_of(h, rq.rows.data(), (int)rq.rows.size(), &n);
      rq.rows.resize(n);
      rq.state = 2;
    } else {
      rq.err = g_err;
      rq.state = 3;
    }
    rq.t_done = tnow;
    s->used -= rq.footprint;  // the arena went back to the pool with fence rq.seq
    rq.h.reset();
    s->completed++;
  }
  s->draining.swap(keep);
}

// A request that cannot be admitted (its arena or uploads failed): it fails alone.
void server_reject(pa_server* s, Request& rq, int rc) {
  rq.rc = rc;
  rq.err = g_err;
  rq.state = 3;
  rq.t_done = now_s();
  if (rq.h && rq.h->arena && rq.h->pooled) s->mem.free_dev_later(rq.h->arena);
  rq.h.reset();
  s->completed++;
}

// Admit queued runs while they fit; returns whether any were admitted.  The initial
// populations of the newcomers are launched on the server stream.
int server_admit(pa_server* s, bool* admitted) {
  std::vector<RunHost*> fresh;
  *admitted = false;
  while (!s->queue.empty() && (int)s->resident.size() < s->max_runs) {
    Request& rq = s->reqs[s->queue.front()];
    if (s->used + rq.footprint > s->budget) break;
    const int rc = materialize(*rq.h, &s->mem);
    if (rc == kWaitForMemory) break;  // pool fragments: wait for retirements
    s->queue.pop_front();
    if (rc != PA_OK) { server_reject(s, rq, rc); continue; }
    s->used += rq.footprint;
    rq.state = 1;
    rq.steps_left = rq.h->n_steps;
    rq.t_admit = now_s();
    s->resident.push_back(rq.ticket);
    fresh.push_back(rq.h.get());
    s->dirty = true;
  }
  if (fresh.empty()) return PA_OK;
  *admitted = true;
  Batch init;
  init.stream = s->batch.stream;  // borrowed
  init.mem = &s->mem;
  init.no_side = true;
  init.runs = fresh;
  int rc = init.build_tables();
  if (!rc) rc = init.launch_init();
  init.stream = nullptr;
  return rc;
}

// Step records and state of a run whose last step has just been enqueued, copied into
// pinned memory behind it; its arena returns to the pool behind the same fence.
int server_drain(pa_server* s, Request& rq) {
  RunHost& h = *rq.h;
  const size_t rec = (h.n_steps + 1) * sizeof(StepOut);
  rq.pin = s->mem.alloc_pin(rec + sizeof(RunState));
  if (!rq.pin) return fail(PA_ERR_HIP, "pinned staging memory for a finished run");
  HIPCHECK(hipMemcpyAsync(rq.pin, h.desc.out, rec, hipMemcpyDeviceToHost, s->batch.stream));
  HIPCHECK(hipMemcpyAsync((uint8_t*)rq.pin + rec, h.desc.st, sizeof(RunState),
                          hipMemcpyDeviceToHost, s->batch.stream));
  s->mem.free_dev_later(h.arena);
  h.arena = nullptr;
  rq.state = 4;
  return PA_OK;
}

}  // namespace

extern "C" {

int pa_server_tick(pa_server* s, int* completed) {
  if (!s) return fail(PA_ERR_INVALID, "null server");
  int rc;
  // At most two ticks in flight: the device always has the next tick queued, and the host
  // sees arrivals and schedules within a tick of the device instead of queueing whole
  // schedules ahead of it.  This waits for the tick before last, never for the last one.
  if (s->ticks >= 2)
    while (s->mem.done() < s->tick_seq[s->ticks & 1]) HIPCHECK(s->mem.wait_oldest());
  server_collect(s);
  bool admitted = false;
  if ((rc = server_admit(s, &admitted))) return rc;
  if (s->dirty) {
    s->batch.runs.clear();
    for (int64_t t : s->resident) s->batch.runs.push_back(s->reqs[t].h.get());
    if ((rc = s->batch.build_tables())) return rc;
    s->dirty = false;
  }
  bool enqueued = admitted;
  if (!s->resident.empty()) {
    // Equal device time per scheduled run (processor sharing): run i earns c_max / c_i
    // steps of credit per tick, c = EqSw * N * Rinit its work per step, and runs the whole
    // part of its credit (at most `chunk`); it pauses on the device after its quota while
    // the tick replays the step max_i k_i times.  Cheap runs therefore finish in few ticks
    // instead of advancing in lockstep with the expensive ones.
    //
    // PA_SERVE_SHORTEST shares the tick only among the runs with the least remaining work
    // (steps left x work per step), taken in that order until their sweep units (workgroups)
    // reach four per CU, so the device stays full while short requests overtake long
    // ones; the others pause for the tick.
    const size_t n = s->resident.size();
    auto cost = [&](int64_t t) {
      const RunHost& h = *s->reqs[t].h;
      return (double)h.p.eqsweeps * h.N * h.R0 + 1.0;
    };
    std::vector<char> on(n, 1);
    if (s->policy == PA_SERVE_SHORTEST) {
      std::vector<size_t> order(n);
      for (size_t i = 0; i < n; ++i) order[i] = i;
      auto rem = [&](size_t i) {
        const int64_t t = s->resident[i];
        return cost(t) * s->reqs[t].steps_left;
      };
      std::stable_sort(order.begin(), order.end(),
                       [&](size_t a, size_t b) { return rem(a) < rem(b); });
      const int64_t fill = 4 * (int64_t)device_cus();
      int64_t units = 0;
      std::fill(on.begin(), on.end(), 0);
      for (size_t j = 0; j < n && units < fill; ++j) {
        const RunHost& h = *s->reqs[s->resident[order[j]]].h;
        on[order[j]] = 1;
        units += (int64_t)h.cap_units * (h.p.coding == PA_SSC ? 1 : h.w / h.sw);
      }
    }
    double cmax = 0;
    for (size_t i = 0; i < n; ++i)
      if (on[i]) cmax = std::max(cmax, cost(s->resident[i]));
    s->pause.resize(n);
    s->quota.resize(n);
    int K = 0;
    for (size_t i = 0; i < n; ++i) {
      Request& rq = s->reqs[s->resident[i]];
      if (on[i]) rq.credit = std::min(rq.credit + cmax / cost(rq.ticket), (double)s->chunk);
      const int k = std::min((int)rq.credit, rq.steps_left);
      rq.credit -= k;
      s->quota[i] = k;
      s->pause[i] = rq.h->n_steps - rq.steps_left + k;  // k = 0: paused for this tick
      K = std::max(K, k);
    }
    HIPCHECK(s->mem.upload(s->batch.pause_d, s->pause.data(), n * sizeof(int32_t)));
    HIPCHECK(pa_launch_resume(s->batch.stream, s->batch.desc_d, (int)n, s->batch.pause_d));
    if ((rc = s->batch.run_steps(K))) return rc;
    enqueued = true;
    // runs whose last step is now enqueued leave the batch; their results follow the step
    std::vector<int64_t> keep;
    for (size_t i = 0; i < n; ++i) {
      const int64_t t = s->resident[i];
      Request& rq = s->reqs[t];
      rq.steps_left -= s->quota[i];
      if (rq.steps_left > 0) { keep.push_back(t); continue; }
      if ((rc = server_drain(s, rq))) return rc;
      s->draining.push_back(t);
      s->dirty = true;
    }
    s->resident.swap(keep);
  }
  if (enqueued || s->mem.has_pending()) {
    int64_t seq = 0;
    HIPCHECK(s->mem.fence(&seq));
    for (int64_t t : s->draining)
      if (s->reqs[t].seq == 0) s->reqs[t].seq = seq;
    if (enqueued) {
      s->tick_seq[s->ticks & 1] = seq;
      ++s->ticks;
    }
  }
  if (!enqueued && !s->mem.idle()) {
    // nothing to enqueue (no run resident, none admissible): wait for the device's oldest
    // fence instead of spinning; it returns finished runs and their memory
    HIPCHECK(s->mem.wait_oldest());
  }
  server_collect(s);
  if (completed) *completed = s->completed;
  return PA_OK;
}

int pa_server_poll(pa_server* s, int64_t ticket, int* done) {
  if (!s || !done) return fail(PA_ERR_INVALID, "null argument");
  server_collect(s);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(PA_ERR_INVALID, "unknown ticket");
  *done = it->second.state == 2 || it->second.state == 3 ? 1 : 0;
  if (it->second.state == 3) return fail(it->second.rc, it->second.err);
  return PA_OK;
}

int pa_server_results(pa_server* s, int64_t ticket, pa_row* rows, int max_rows, int* nrows) {
  if (!s || !rows || !nrows) return fail(PA_ERR_INVALID, "null argument");
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return fail(PA_ERR_INVALID, "unknown ticket");
  if (it->second.state == 3) return fail(it->second.rc, it->second.err);
  if (it->second.state != 2) return fail(PA_ERR_STATE, "request not finished");
  const int n = std::min((int)it->second.rows.size(), max_rows);
  memcpy(rows, it->second.rows.data(), n * sizeof(pa_row));
  *nrows = n;
  return PA_OK;
}

int pa_server_latency(pa_server* s, int64_t ticket, double* submit_to_done_s, double* admit_to_done_s) {
  if (!s) return fail(PA_ERR_INVALID, "null server");
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second.state != 2) return fail(PA_ERR_STATE, "request not finished");
  if (submit_to_done_s) *submit_to_done_s = it->second.t_done - it->second.t_submit;
  if (admit_to_done_s) *admit_to_done_s = it->second.t_done - it->second.t_admit;
  return PA_OK;
}

int pa_server_stats(pa_server* s, int* queued, int* resident, size_t* hbm_used) {
  if (!s) return fail(PA_ERR_INVALID, "null server");
  server_collect(s);
  if (queued) *queued = (int)s->queue.size();
  if (resident) *resident = (int)(s->resident.size() + s->draining.size());
  if (hbm_used) *hbm_used = s->used;
  return PA_OK;
}

int pa_server_destroy(pa_server* s) {
  if (!s) return PA_OK;
  (void)hipStreamSynchronize(s->batch.stream);
  s->batch.runs.clear();
  delete s;
  return PA_OK;
}

}  // extern "C"
