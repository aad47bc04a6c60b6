/*
 * paising.h - C ABI of the MI355X population-annealing engine (libpaising.so).
 *
 * The reference (LevBarash/PAising) has no library API: its boundary is the process
 * (argv in; files PA_results_run_<r>.dat in data<SSC|MSC>_L%d_R%d_EqSw%d_dB%f, stdout and exit code
 * out; PAisingSSC.cu:381-806, PAisingMSC.cu:546-957).  The drop-in executables PAisingSSC
 * and PAisingMSC (paising_amd/csrc/pa_cli.cpp) keep that boundary.  This header is the
 * C ABI those executables, the batcher, the server and foreign-language bindings call; each
 * entry point names the reference code it replaces.
 *
 * Conventions: plain C types, caller-owned opaque handles, integer error codes (PA_OK = 0),
 * no exceptions across the boundary, one host thread per handle, work ordered on the
 * handle's HIP stream.  The last error message of the calling thread is pa_last_error().
 */
#ifndef PAISING_H
#define PAISING_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pa_status {
  PA_OK = 0,
  PA_ERR_INVALID = 1,     /* bad parameters                                         */
  PA_ERR_HIP = 2,         /* HIP runtime error (the reference exits with the code)  */
  PA_ERR_NBMAX = 3,       /* more than nBmax-2 steps (PAisingSSC.cu:634-641)        */
  PA_ERR_CAPACITY = 4,    /* population outgrew its buffers (handled internally)    */
  PA_ERR_MEMORY = 5,      /* request does not fit the HBM budget                    */
  PA_ERR_STATE = 6        /* call not valid in the handle's current state           */
};

enum pa_coding { PA_SSC = 0, PA_MSC = 1 };

/* uniform-double convention of the RNG stream (one 32-bit word per double):
 * PA_UNIFORM_CUDA101 = (v+1)/2^32 reproduces the CUDA >= 10.1 reference outputs,
 * PA_UNIFORM_CUDA100 = (v+1/2)/2^32 the CUDA <= 10.0 ones. */
enum pa_uniform { PA_UNIFORM_CUDA101 = 0, PA_UNIFORM_CUDA100 = 1 };
/* Server tick policies (pa_server_set_policy). */
enum pa_serve_policy { PA_SERVE_SHARE = 0, PA_SERVE_SHORTEST = 1 };

/* Run parameters: the reference's command-line options (PAisingSSC.cu:386-400) plus the
 * compile-time L and MSbits (PAisingSSC.cu:27, PAisingMSC.cu:34), made runtime. */
typedef struct pa_params {
  int32_t L;                 /* linear lattice size, even, >= 4 (reference: 64)          */
  int32_t coding;            /* PA_SSC or PA_MSC                                          */
  int32_t msbits;            /* MSC replicas per word: 32 or 64 (reference: 32)           */
  int32_t Rinit;             /* -R initial population (20000)                             */
  uint32_t eqsweeps;         /* -t sweeps per temperature (100)                           */
  double dB;                 /* -d inverse-temperature step (0.005)                       */
  double Binit;              /* initial beta (0)                                          */
  double Bfin;               /* -f final beta (1)                                         */
  uint64_t seed;             /* -s seed, as the reference passes it to curand_init         */
  uint64_t initial_sequence; /* Philox subsequence counter at the start of the run (0)    */
  int32_t uniform;           /* pa_uniform                                                */
  int32_t max_steps;         /* 0 = whole schedule; otherwise stop after this many steps  */
  int32_t trace;             /* 1 = record per-step population hashes (parity checks)     */
  int32_t slice_bits;        /* MSC sweep slice width 8/16/32/64, 0 = autotuned; SSC:     */
                             /* 1 = site-per-thread kernels instead of packed-byte ones,  */
                             /* 2 = the general packed-byte kernel also where L/2 x L    */
                             /* is a multiple of 1024 (L % 8 == 0 lattices)               */
  int32_t threads;           /* sweep workgroup size 128/256/512 (MSC also 1024; packed   */
                             /* SSC any multiple of 64 up to 1024), 0 = autotuned         */
  int32_t balance;           /* MSC makespan-balanced sweeps: 0 auto, 1 on, -1 off        */
  int32_t adaptive;          /* 1 = adaptive beta step, the reference's AdaptiveStep build  */
                             /* (PAisingSSC.cu:23,644-660): plain runs only               */
  int32_t schedule_len;      /* > 0: run the explicit schedule B[0..len-1] instead of the  */
                             /* dB ladder (replay of run 0's steps, PAisingSSC.cu:662)    */
  double min_overlap;        /* adaptive: MinOverlap, 0 = 0.85 (PAisingSSC.cu:41)          */
  double max_overlap;        /* adaptive: MaxOverlap, 0 = 0.87 (PAisingSSC.cu:42)          */
  const double* schedule;    /* explicit schedule (B[0] = Binit), copied at creation      */
  int32_t mhr;               /* 1 = accumulate the energy histogram over all steps for    */
                             /* multi-histogram reweighting (MHR build, PAisingSSC.cu:22, */
                             /* 294-304,496-501,608-611)                                  */
  int32_t pad1;
} pa_params;

/* One row of PA_results_run_<r>.dat (PAisingSSC.cu:679-690) and the step internals. */
typedef struct pa_row {
  double B, E, C, M, M2, M4, BF, S; /* as printed: E/N, C/N, M/N, M2/N^2, M4/N^4, BF/N, S/N */
  int32_t R;
  int32_t Rprev;
  double lnQ;
  double q;
  double sums[5];            /* sum E, E^2, |M|, M^2, M^4 over the population        */
  int64_t sumE, sumM;
  uint64_t roff_hash, spin_hash, IS_after;
} pa_row;

typedef struct pa_run pa_run;
typedef struct pa_batch pa_batch;
typedef struct pa_server pa_server;

const char* pa_last_error(void);
int pa_version(void);
int pa_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes);

/* Streaming device-to-device copy of `bytes` (read + write), `iters` times: achieved GB/s
 * (counts read and written bytes).  No reference counterpart: the HBM reference point for
 * the resampling kernels and for scaling the FETCH_SIZE / WRITE_SIZE counters. */
int pa_bandwidth_probe(size_t bytes, int iters, double* gbps);

/* Number of beta steps of the reference schedule B_i = B_{i-1} + dB while B_i <= Bfin
 * (PAisingSSC.cu:508-510,661). */
int pa_schedule_steps(double Binit, double dB, double Bfin);

/* ---- single run: replaces main()'s run-loop body, PAisingSSC.cu:481-692 ---- */
int pa_run_create(const pa_params* p, pa_run** out);
/* Enqueue up to nsteps beta steps (negative = all remaining); asynchronous. */
int pa_run_step(pa_run* r, int nsteps);
int pa_run_sync(pa_run* r);
int pa_run_steps_done(pa_run* r, int* done, int* total);
/* Rows 0..n of the results file (host thermodynamics, PAisingSSC.cu:620-630,679-690). */
int pa_run_results(pa_run* r, pa_row* rows, int max_rows, int* nrows);
/* Philox subsequence counter after the last completed step (continues across -M runs,
 * PAisingSSC.cu:418). */
int pa_run_final_sequence(pa_run* r, uint64_t* is);
/* Copy the current population to host memory: SSC R x N int8 (gA then gB per replica),
 * MSC ceil(R/w) x N words; returns bytes needed in *bytes when dst is NULL. */
int pa_run_population(pa_run* r, void* dst, size_t* bytes);
/* Per-member arrays of the current population: 0 = IE, 1 = M (int32), 2 = Roff, 3 = prefix
 * (uint32, of the last resampling), 4 = MSC active masks (uint64 per group). */
int pa_run_array(pa_run* r, int which, void* dst, size_t* bytes);
/* Elapsed device time of the sweep kernels and of the whole run so far (ms). */
int pa_run_timing(pa_run* r, double* sweep_ms, double* total_ms);
/* Device time of the resampling kernels and of the sweep kernels summed over the timed
 * steps, and their count (valid after pa_run_sync).  The reference times only the whole
 * program (PAisingSSC.cu:445, 801-806). */
int pa_run_kernel_times(pa_run* r, double* resample_ms, double* sweep_ms, int* steps);
/* Time each step's resampling and sweep launches with HIP events on the run's stream
 * (1 = on); the events are read at the next pa_run_sync, so timing adds no host sync. */
int pa_run_set_timing(pa_run* r, int on);
/* Replay steps from a captured hipGraph (1 = on). */
int pa_run_set_graph(pa_run* r, int on);
int pa_run_destroy(pa_run* r);

/* Autotuner (no reference counterpart: the reference fixes 128 threads per replica,
 * PAisingSSC.cu:62).  Times every valid sweep variant (MSC slice width, workgroup size,
 * makespan balancing) on a two-step trial of *p, writes the fastest into *tuned and caches it
 * for this process: later runs of the same shape (coding, w, L, log2 Rinit) that leave
 * slice_bits/threads/balance at 0 use it.  Results never depend on the variant. */
int pa_autotune(const pa_params* p, pa_params* tuned, double* best_sweep_ms);
int pa_autotune_clear(void);

/* Whole run in one call: create, all steps, results (a reference run, PAisingSSC.cu:481-692). */
int pa_run_all(const pa_params* p, pa_row* rows, int max_rows, int* nrows, uint64_t* final_is);
/* pa_run_all that also returns the run's energy histogram (p->mhr = 1): hist[k] counts the
 * replicas with energy 4k - 2N summed over the initial population and every step
 * (ShistE, PAisingSSC.cu:496-501,608-611); nbins must be at least N + 1. */
int pa_run_all_hist(const pa_params* p, pa_row* rows, int max_rows, int* nrows,
                    uint64_t* final_is, int32_t* hist, int nbins);
/* The energy histogram of a run created with p->mhr = 1 (after pa_run_sync). */
int pa_run_histogram(pa_run* r, int32_t* hist, int nbins);
/* Multi-histogram reweighting of one run (PAisingSSC.cu:695-764, host FP64): nB rows with
 * inverse temperatures B, populations R, dimensionless free energies BF (unnormalised:
 * BF[0] = -N ln 2, BF[i] = BF[0] - sum of lnQ; updated in place by every iteration, as the
 * reference does), the histogram of N + 1 bins, niter iterations.  out receives nB rows of
 * 1 + 3 niter doubles: B, then per iteration E/N, C/N, BF/N (PA_MHR_results_run_<r>.dat). */
int pa_mhr_analysis(int N, int nB, const double* B, double* BF, const int32_t* R,
                    const int32_t* hist, int niter, double* out);

/* ---- single stages of a beta step on caller data (host arrays in and out) ----
 * Each call builds a scratch device run (it allocates), fills its state from the
 * arguments, launches the kernels of that part of a step -- the same launches a run makes --
 * and copies the results back.  p supplies L, coding, msbits, seed, the uniform mapping,
 * EqSw, Rinit (copy counts) and the sweep variant; the other schedule fields are ignored.
 * Layouts as pa_run_population: SSC R x N int8, MSC ceil(R/w) x N words; per-member arrays
 * have R (SSC) or ceil(R/w)*w (MSC) entries; `active` holds one byte per MSC lane (NULL:
 * lanes below R). */
/* checkKerALL + energyKer (PAisingSSC.cu:159-215, PAisingMSC.cu:190-317): EqSw sweeps of R
 * replicas at inverse temperature beta, sweep streams starting at subsequence `is`
 * (unit B uses is + 128 B + t), then energies and magnetisations. */
int pa_stage_sweep(const pa_params* p, int R, uint64_t is, double beta, void* pop,
                   int32_t* IE, int32_t* M);
/* pa_stage_sweep on device memory, allocating nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  pop (the population, layout as above), IE and M
 * (ceil(R/w)*w or R int32 each) are device pointers owned by the caller; the work is enqueued
 * on `stream` (a hipStream_t, NULL = the null stream) and the call returns without waiting.
 * The run's descriptor, state and thresholds go to a fixed 4096-byte device array of the
 * library, one per device, kept until the process ends; makespan balancing is not used (it
 * needs flags per unit), which leaves the results unchanged.  Calls that overlap on one
 * device race on that array: order them on one stream or wait between them. */
int pa_stage_sweep_noalloc(const pa_params* p, int R, uint64_t is, double beta, void* pop,
                           int32_t* IE, int32_t* M, void* stream);
/* energyKer alone (PAisingSSC.cu:195-215, PAisingMSC.cu:268-317). */
int pa_stage_energy(const pa_params* p, int R, const void* pop, int32_t* IE, int32_t* M);
/* QKer, lnQ, CalcTauKer and CalcParSum (PAisingSSC.cu:217-270,533; PAisingMSC.cu:319-395,
 * 696): q and lnQ of parents with energies IE at increment dB around Emean, copy counts
 * drawn from subsequences is + n, their exclusive prefix and R_new. */
int pa_stage_copy_counts(const pa_params* p, int R, const int32_t* IE, const uint8_t* active,
                         double dB, double Emean, uint64_t is, double* q, double* lnQ,
                         uint32_t* Roff, uint32_t* prefix, int* Rnew);
/* resampleKer (PAisingSSC.cu:272-280, PAisingMSC.cu:397-435): the children of R parents
 * with copy counts Roff (MSC: child n -> group n mod rg_new, lane n div rg_new). */
int pa_stage_resample(const pa_params* p, int R, const void* pop, const uint8_t* active,
                      const uint32_t* Roff, void* pop_new, uint8_t* active_new, int* Rnew);
/* CalcAverages (PAisingSSC.cu:282-292, PAisingMSC.cu:437-450): sums of E, E^2, |M|, M^2, M^4
 * over the active members in the device's fixed reduction order, and the integer sums. */
int pa_stage_averages(const pa_params* p, int R, const int32_t* IE, const int32_t* M,
                      const uint8_t* active, double sums[5], int64_t* sumE, int64_t* sumM);

/* ---- batched runs (no reference counterpart: many independent runs per launch) ---- */
int pa_batch_create(const pa_params* p, int nruns, int use_graph, pa_batch** out);
int pa_batch_run(pa_batch* b);            /* all runs to completion */
int pa_batch_results(pa_batch* b, int run, pa_row* rows, int max_rows, int* nrows);
/* Device time of the last pa_batch_run (ms). */
int pa_batch_time(pa_batch* b, double* total_ms);
int pa_batch_destroy(pa_batch* b);

/* ---- serving (no reference counterpart): continuous batching under an HBM budget ---- */
int pa_server_create(size_t hbm_budget_bytes, int max_runs, pa_server** out);
int pa_server_submit(pa_server* s, const pa_params* p, int64_t* ticket);
/* One tick: pick up the results of runs whose last step has completed, admit queued
 * requests while their footprint fits the budget (a request whose device state cannot be
 * created fails alone), and enqueue the next steps of the scheduled resident runs (see
 * pa_server_set_chunk / pa_server_set_policy).  Asynchronous: at most two ticks are in
 * flight, and a tick never waits for the device to drain.  *completed = requests finished
 * (or failed) so far. */
int pa_server_tick(pa_server* s, int* completed);
/* *done = 1 when the ticket has finished or failed, 0 when queued or running; a failed
 * request returns its error code. */
int pa_server_poll(pa_server* s, int64_t ticket, int* done);
int pa_server_results(pa_server* s, int64_t ticket, pa_row* rows, int max_rows, int* nrows);
int pa_server_stats(pa_server* s, int* queued, int* resident, size_t* hbm_used);
/* Scheduling of a tick: every resident run gets about the same device time, i.e. a step
 * quota inversely proportional to its work per step (EqSw * N * Rinit), at least 1 and at
 * most steps_per_tick (default 32); use_graph: ticks replay a captured hipGraph. */
int pa_server_set_chunk(pa_server* s, int steps_per_tick, int use_graph);
/* Which resident runs share a tick.  PA_SERVE_SHARE: all of them (processor sharing).
 * PA_SERVE_SHORTEST (default): the runs with the least remaining work (steps left x work per
 * step), in that order, until their workgroups reach four per CU; the rest pause for the tick.
 * Results do not depend on the policy.  No counterpart in the reference (serving only). */
int pa_server_set_policy(pa_server* s, int policy);
/* Device memory of the server (no reference counterpart).  Run arenas come from slabs of
 * device memory that only grow (at least 4 GiB each, within the budget) and are reused after
 * runs retire, so admissions and retirements neither allocate nor free device memory once
 * the pool holds the resident load.  pa_server_reserve grows the arena pool to `bytes` (at
 * most the budget) up front; pa_server_memory reports the device bytes the server holds
 * (arena and table slabs) and how many device allocations it has made so far. */
int pa_server_reserve(pa_server* s, size_t bytes);
int pa_server_memory(pa_server* s, size_t* reserved, int* device_allocations);
/* Seconds from submission (and from admission) to completion of a finished request. */
int pa_server_latency(pa_server* s, int64_t ticket, double* submit_to_done_s,
                      double* admit_to_done_s);
int pa_server_destroy(pa_server* s);

/* Footprint in bytes of one run's device state (admission control). */
size_t pa_run_footprint(const pa_params* p);

#ifdef __cplusplus
}
#endif
#endif
