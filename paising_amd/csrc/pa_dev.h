// Device-resident run descriptors shared by the host engine and the kernels.
//
// One PA run (one reference program invocation's run loop, PAisingSSC.cu:481-666) keeps its
// whole population in HBM for the whole schedule.  Everything a step needs that depends on
// the step index (beta, dB, acceptance thresholds) is precomputed by the host into per-step
// tables, and everything a step produces (R, q, lnQ, sums) is written to per-step records,
// so a step is a fixed sequence of kernel launches with no host round trip.  Several runs
// (a batch) are processed by the same launches: every kernel maps its blocks to runs
// through a per-kernel prefix table.
#ifndef PA_DEV_H
#define PA_DEV_H

#include <stdint.h>

namespace pa {

enum Coding : int32_t { kSSC = 0, kMSC = 1 };

// MSC resampling of large populations by windows of new groups (k_resample_win): new groups
// per window (one thread each), parent rows staged per 128-byte site chunk, and the
// workgroup's dynamic LDS (the rows, a zero row, the row -> parent group table).
constexpr int kWinJ = 256;
constexpr int kWinSlots = 511;
constexpr size_t kWinLds = (size_t)(kWinSlots + 1) * 128 + (size_t)kWinSlots * 4;

// Mutable state of one run, advanced by the kernels.
struct RunState {
  int32_t step;      // index i of the step in progress (1-based), == rows done + 1
  int32_t R;         // population size R_{i-1} at the start of the step
  int32_t Rnew;      // population size R_i after resampling
  int32_t buf;       // spin buffer holding the current population
  uint64_t IS;       // Philox subsequence counter ("initial_sequence", PAisingSSC.cu:418)
  double Emean;      // population mean energy of the current population (0 before step 1)
  double lnQ;        // lnQ_i of the step in progress
  int32_t overflow;  // set when R_i exceeds the buffer capacity; the run then stops
  int32_t active;    // 0 once the run has finished its schedule
  int32_t err;       // device-side error bits (1: balanced-sweep hand-over wait timed out)
  int32_t pause;     // serving: the run stops (active = 0) after this step; 0 = never
  int32_t qdone;     // k_qgroup blocks finished in the current launch (last-block election)
  int32_t avdone;    // k_avg blocks finished in the current launch
};

// Per-step record written by the finalize kernel (row i of the results file + internals).
struct StepOut {
  double q;          // sum of exp(-dB (E - Emean)) (QKer)
  double lnQ;
  double sums[5];    // sum E, E^2, |M|, M^2, M^4 over the new population
  int64_t sumE;      // integer sums (parity checks)
  int64_t sumM;
  int32_t R;         // R_i
  int32_t Rprev;     // R_{i-1}
  uint64_t roff_hash;
  uint64_t spin_hash;
  uint64_t IS_after;
};

struct RunDesc {
  int32_t L, Ld2, N, Nh;
  int32_t coding, w;        // w: replicas per word (MSC), 1 for SSC
  int32_t Rinit, n_steps;
  int32_t cap_units;        // replicas (SSC) or groups (MSC) per spin buffer
  int32_t umap;             // uniform-double mapping (kUniformOpenClosed / kUniformMidpoint)
  uint32_t eqsweeps;
  int32_t trace;            // compute per-step hashes
  int32_t q128, r128;       // 128 / Ld2, 128 % Ld2 (site-walk increments)
  uint32_t ld2_magic;       // floor(2^32 / Ld2) + 1: idx / Ld2 == mulhi(idx, magic) for idx < 2^17
  int32_t rs_lp;            // MSC resampling: site chunks per 128-byte line (same XCD)
  int32_t ssc_p;            // packed SSC sweeps (k_sweep_ssc_p): parts per stream quad
  int32_t pad0;
  uint64_t seed;
  void* pop[2];             // spin buffers: SSC int8 [cap][N], MSC word [cap][N]
  uint64_t* amask[2];       // MSC active-lane bit masks per group (ping-pong with pop)
  int32_t* IE;              // energies    [cap_pop]
  int32_t* M;               // magnetisations [cap_pop]
  uint32_t* Roff;           // copy counts [cap_pop]
  uint32_t* pre;            // exclusive prefix of Roff [cap_pop]
  int32_t* par;             // parent of every child of the last resampling [cap_pop]
  uint32_t* flags;          // balanced sweeps: hand-over flag per (group, slice) [cap_units*8]
  double* gsum;             // MSC per-group lane sums of the q terms [cap_units]
  int32_t* hist;            // MHR energy histogram [N + 1] (null unless requested)
  int32_t rs_sites;         // MSC resampling: site words per workgroup chunk
  int32_t rs_split;         // MSC resampling: workgroups sharing one site chunk (new groups split)
  uint32_t* cblk;           // per-1024 chunk sums of Roff
  double* qblk;             // per-1024 chunk partial q
  double* avblk;            // per (lane, chunk) partial averages, 5 each
  int64_t* iblk;            // per chunk integer sums (E, M)
  uint64_t* hblk;           // per chunk hash partials
  const double* beta;       // [n_steps + 1]
  const double* dbeta;      // [n_steps + 1]  B[i] - B[i-1]
  const uint64_t* thr;      // [n_steps + 1][3]: SSC thresholds for dE = 0, 4, 8; MSC T4, T8
  StepOut* out;             // [n_steps + 1]
  RunState* st;
};

// Work-split prefix tables: for kernel kind k, runs[r] owns blocks
// [offs[k*(nruns+1) + r], offs[k*(nruns+1) + r + 1]).
enum KernelKind : int32_t {
  kKUnits = 0,        // one block per unit (replica / group) of capacity
  kKChunks = 1,       // one block per 1024 population members (replicas / lanes)
  kKGroupChunks = 2,  // q partials: 1024 replicas (SSC) or 1024 groups (MSC)
  kKAvg = 3,          // averages: 1024 replicas (SSC) or (1024 groups, lane) (MSC)
  kKResample = 4,     // resampling: one block per child replica (SSC) / site chunk (MSC)
  kKSweep = 5,        // packed SSC sweeps: one block per RPW replicas (k_sweep_ssc_p)
  kKindCount = 6
};

}  // namespace pa

#endif
