/*
 * CPU oracle for the population-annealing (PA) engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain host C++/OpenMP restatement of the
 * reference programs PAisingSSC.cu and PAisingMSC.cu (LevBarash/PAising), used as
 * the checker for the HIP engine.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product (paising_amd/) never
 * links, imports or calls anything here.
 *
 * Every routine cites the reference lines it restates (PAisingSSC.cu = SSC,
 * PAisingMSC.cu = MSC).  Pinning of each routine is documented in pa_oracle.cpp.
 */
#ifndef PA_ORACLE_H
#define PA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORA_SSC = 0, ORA_MSC = 1 };

/* Uniform-double mappings of curand_uniform_double (candidates from named open sources).
 * The fixtures select a single 32-bit word per double (rung lnQ_1 of the SSC fixture); the
 * two-word candidates 0..5 are kept so the test can show that they are rejected. */
enum {
  ORA_U_ROCRAND_UNIFORM = 0, /* rocrand_uniform.h:102-109  z = x | (y>>11)<<32, u = (z+1)/2^53 */
  ORA_U_XOR21_P1 = 1,        /* rocrand_normal.h:81-83     z = x ^ y<<21,       u = (z+1)/2^53 */
  ORA_U_XOR21_HALF = 2,      /* z = x ^ y<<21, u = z/2^53 + 2^-54 (Random123 open-open offset) */
  ORA_U_XOR21_SWAP = 3,      /* z = y ^ x<<21, u = (z+1)/2^53 */
  ORA_U_U64_XHI = 4,         /* rocrand_uniform.h:111-117 on v = x<<32|y */
  ORA_U_U64_YHI = 5,         /* rocrand_uniform.h:111-117 on v = y<<32|x */
  ORA_U32_OPEN_CLOSED = 6,   /* rocrand_uniform.h:97-100: u = (v+1)/2^32, in (0,1]          */
  ORA_U32_MIDPOINT = 7,      /* Random123 u01fixedpt open-open convention: u = (v+1/2)/2^32 */
  ORA_U32_CLOSED_OPEN = 8    /* Random123 u01fixedpt closed-open convention: u = v/2^32     */
};

typedef struct {
  int32_t L;            /* linear size, even                                  */
  int32_t coding;       /* ORA_SSC or ORA_MSC                                 */
  int32_t msbits;       /* replicas per word for MSC: 32 or 64                */
  int32_t Rinit;        /* initial population                                 */
  uint32_t eqsweeps;    /* sweeps per temperature step (-t)                   */
  double dB;            /* inverse-temperature step (-d)                      */
  double Binit;         /* initial beta (0 in the reference)                  */
  double Bfin;          /* final beta (-f)                                    */
  int32_t runs;         /* independent runs (-M)                              */
  uint64_t seed;        /* RNG seed (-s), already sign-extended               */
  uint64_t initial_sequence; /* starting Philox subsequence counter           */
  int32_t umap;         /* ORA_U_* mapping                                    */
  int32_t max_steps;    /* stop after this many beta steps (0 = full schedule) */
  int32_t threads;      /* OpenMP threads (0 = runtime default)               */
  int32_t adaptive;     /* 1 = adaptive beta step (#define AdaptiveStep, SSC:23) */
  int32_t pad0;
  double min_overlap;   /* MinOverlap (SSC:41)                                */
  double max_overlap;   /* MaxOverlap (SSC:42)                                */
  int32_t mhr;          /* 1 = energy histograms + MHR analysis (#define MHR, SSC:22) */
  int32_t mhr_iter;     /* MHR_Niter (SSC:53), 0 = 1                          */
} ora_params;

/* One row of PA_results_run_r.dat (SSC:679-690) plus the per-step internals the
 * GPU engine is compared against.  Row 0 is the beta=Binit row. */
typedef struct {
  double B, E, C, M, M2, M4, BF, S;
  int32_t R;
  int32_t Rprev;
  double lnQ;
  double q;             /* partition-function ratio sum (QKer result)         */
  double sums[5];       /* sum E, E^2, |M|, M^2, M^4 over the population       */
  int64_t sumE_int;     /* integer sum of replica energies                    */
  int64_t sumM_int;     /* integer sum of replica magnetisations              */
  uint64_t roff_hash;   /* hash of the copy-count array in population order   */
  uint64_t spin_hash;   /* hash of the spin configuration after the sweeps    */
  uint64_t IS_after;    /* Philox subsequence counter after the step          */
} ora_row;

/* ora_run plus the MHR build's outputs: hist (runs * (N+1) ints) receives each run's energy
 * histogram, mhr (runs * max_rows * (1 + 3 mhr_iter) doubles) each run's
 * PA_MHR_results_run_r.dat rows (B, then E/N, C/N, BF/N per iteration).  Either may be
 * null. */
int ora_run_ex(const ora_params* p, ora_row* rows, int max_rows, int32_t* nrows,
               double* elapsed_s, double* total_pop, int32_t* hist, double* mhr);
/* The MHR analysis of SSC:705-764 on given arrays (BF updated in place). */
void ora_mhr_analysis(int N, int nB, const double* B, double* BF, const int32_t* R,
                      const int32_t* hist, int niter, double* out);

/* Full PA run(s).  rows has room for runs*max_rows rows; run r uses rows
 * [r*max_rows, r*max_rows + nrows[r]).  Returns 0, or 1 when the step count
 * reaches nBmax (SSC:634-641), or -1 on bad parameters. */
int ora_run(const ora_params* p, ora_row* rows, int max_rows, int32_t* nrows,
            double* elapsed_s, double* total_pop);

/* ---- primitives ---- */
void ora_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint32_t ora_stream_word(uint64_t seed, uint64_t subsequence, uint64_t pos);
double ora_uniform(uint32_t x, uint32_t y, int umap);
double ora_exp(double x);
double ora_log(double x);
uint64_t ora_hash_bytes(uint64_t h, const void* data, uint64_t n);
double ora_block_reduce(const double* v, int n); /* reference blockReduceSum tree */

/* ---- stages on a caller-owned population (used by the GPU parity tests) ----
 * SSC spins: R replicas x N int8 (gA[N/2] then gB[N/2]).
 * MSC words: rg groups x N words (gA then gB), uint32 (w=32) or uint64 (w=64),
 * active: rg*w bytes, IE/M: rg*w int32. */
void ora_ssc_init(int L, int R, uint64_t seed, uint64_t IS, int umap, int8_t* spins);
void ora_ssc_sweep(int L, int R, int8_t* spins, uint32_t sweeps, uint64_t seed,
                   uint64_t IS, double beta, int umap);
void ora_ssc_energy(int L, int R, const int8_t* spins, int32_t* IE, int32_t* M);
void ora_msc_init(int L, int w, int rg, int R, uint64_t seed, uint64_t IS, void* words,
                  uint8_t* active);
void ora_msc_sweep(int L, int w, int rg, void* words, uint32_t sweeps, uint64_t seed,
                   uint64_t IS, double beta);
void ora_msc_energy(int L, int w, int rg, const void* words, int32_t* IE, int32_t* M);
/* Copy counts (CalcTauKer, SSC:236-254 / MSC:347-360); returns sum of Roff. */
int64_t ora_ssc_tau(int R, const int32_t* IE, int Rinit, double dB, double lnQ,
                    uint64_t seed, uint64_t IS, int umap, uint32_t* Roff);
int64_t ora_msc_tau(int w, int rg, const uint8_t* active, const int32_t* IE, int Rinit,
                    int R, double dB, double lnQ, uint64_t seed, uint64_t IS, int umap,
                    uint32_t* Roff);
double ora_ssc_q(int R, const int32_t* IE, double dB, double Emean);
double ora_msc_q(int w, int rg, const uint8_t* active, const int32_t* IE, double dB,
                 double Emean);
double ora_lnq(double dB, double Emean, double q, int Rprev);
/* CalcOverlap (SSC:317-334, MSC:477-499): energy-histogram overlap of the current
 * population for a candidate step dB. */
double ora_ssc_overlap(int R, const int32_t* IE, int Rinit, double dB, double Emean);
double ora_msc_overlap(int w, int rg, const uint8_t* active, const int32_t* IE, int Rinit,
                       int R, double dB, double Emean);
void ora_ssc_averages(int R, const int32_t* IE, const int32_t* M, double out[5]);
void ora_msc_averages(int w, int rg, const uint8_t* active, const int32_t* IE,
                      const int32_t* M, double out[5]);
/* Resampling: exclusive prefix then copy/scatter into a new population. */
void ora_ssc_resample(int L, int Rold, const int8_t* spins, const uint32_t* Roff,
                      int8_t* spins_new);
void ora_msc_resample(int L, int w, int rg_old, const void* words, const uint8_t* active,
                      const int32_t* IE, const uint32_t* Roff, int Rnew, void* words_new,
                      uint8_t* active_new, int32_t* IE_new);
/* Boltzmann tables: SSC 10 FP64 entries (SSC:512-516); MSC 2 uint32 (MSC:674-675). */
void ora_ssc_table(double beta, double T[10]);
void ora_msc_table(double beta, uint32_t T[2]);

#ifdef __cplusplus
}
#endif
#endif
