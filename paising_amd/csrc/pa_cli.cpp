// Drop-in executables PAisingSSC (PA_CLI_CODING=0) and PAisingMSC (PA_CLI_CODING=1).
//
// Same options, defaults, directory naming, results-file format, stdout lines and exit
// codes as the reference main() (PAisingSSC.cu:338-806, PAisingMSC.cu:503-957), with the
// compile-time constants made runtime options:
//   -L size   (PAisingSSC.cu:27, default 64)
//   -w bits   (PAisingMSC.cu:34 MSbits, 8, 16, 32 or 64, default 32; MSC only)
//   -u conv   uniform-double convention: 0 = CUDA >= 10.1 (default), 1 = CUDA <= 10.0
//   -a        adaptive inverse-temperature step (the reference's AdaptiveStep build,
//             PAisingSSC.cu:23,40-43,406-410,644-660); later -M runs replay run 0's steps
//   -H        multi-histogram reweighting (the reference's MHR build, PAisingSSC.cu:22,52-54,
//             695-791): PA_MHR_results_run_<r>.dat with MHR_Niter = 1
//   -e        PA_energies_<i>.dat after every step (EnergiesPopStore, PAisingSSC.cu:24,
//             563-578; PAisingMSC.cu:743-757)
//   -l        PA_lattice_<i>.dat after every step of run 0 (LatticePopStore, SSC only,
//             PAisingSSC.cu:25,424-427,580-605)
// The GPU work goes through the C ABI of paising.h (pa_run_all per run).
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <vector>

#include "paising.h"

#ifndef PA_CLI_CODING
#define PA_CLI_CODING 0
#endif

namespace {

// The 4.4BSD getopt(3) (University of California, Berkeley; BSD licence), which the
// reference carries as its own option parser (PAisingSSC.cu:338-365), kept with its
// behaviour because the option quirks and messages are part of the command-line contract:
// prints "illegal option -- c" / "option requires an argument -- c" and returns '?'.
char* opt_arg;
int opt_ind = 1, opt_opt;

int get_opt(int nargc, char* const nargv[], const char* ostr) {
  static const char* place = "";
  const char* oli;
  if (!*place) {
    if (opt_ind >= nargc || *(place = nargv[opt_ind]) != '-') { place = ""; return -1; }
    if (place[1] && *++place == '-') { ++opt_ind; place = ""; return -1; }
  }
  if ((opt_opt = (int)*place++) == (int)':' || !(oli = strchr(ostr, opt_opt))) {
    if (opt_opt == (int)'-') return -1;
    if (!*place) ++opt_ind;
    printf("illegal option -- %c\n", opt_opt);
    return (int)'?';
  }
  if (*++oli != ':') {
    opt_arg = nullptr;
    if (!*place) ++opt_ind;
  } else {
    if (*place) {
      opt_arg = (char*)place;
    } else if (nargc <= ++opt_ind) {
      place = "";
      printf("option requires an argument -- %c\n", opt_opt);
      return (int)'?';
    } else {
      opt_arg = nargv[opt_ind];
    }
    place = "";
    ++opt_ind;
  }
  return opt_opt;
}

// The reference's usage text (PAisingSSC.cu:367-378) word for word for its own options (part
// of the drop-in contract), followed by the options this build adds.
void usage(const char* prog) {
  std::cout << " Usage: " << prog << " [options]\n"
            << " Note: all of the options are optional. Default parameter values are listed in the head of the source code. \n"
            << " Possible command line options are:\n\n"
            << " -R Rinit           ( Rinit           = initial size of population of replicas )\n"
            << " -t EQsweeps        ( EQsweeps        = number of equilibration sweeps )\n"
            << " -d dBinit          ( dBinit          = inverse temperature step )\n"
            << " -f Bfin            ( Bfin            = final value of inverse temperature )\n"
            << " -M runs            ( runs            = number of population annealing algorithm independent runs )\n"
            << " -s RNGseed         ( RNGseed         = seed for random number generation )\n"
            << " -P OutputPrecision ( OutputPrecision = precision (number of digits) of the output )\n"
            << " -o dataDirectory   ( dataDirectory   = data directory name )\n"
            << " -L size            ( size            = linear lattice size, even )\n"
#if PA_CLI_CODING == 1
            << " -w bits            ( bits            = replicas per multi-spin word, 8, 16, 32 or 64 )\n"
#endif
            << " -u convention      ( convention      = uniform-double mapping: 0 CUDA>=10.1, 1 CUDA<=10.0 )\n"
            << " -a                 ( adaptive inverse temperature step, MinOverlap 0.85, MaxOverlap 0.87 )\n"
            << " -H                 ( multi-histogram reweighting analysis, PA_MHR_results_run_<r>.dat )\n"
            << " -e                 ( store the population's energies at each temperature, PA_energies_<i>.dat )\n"
#if PA_CLI_CODING == 0
            << " -l                 ( store the population's lattices at each temperature of run 0, PA_lattice_<i>.dat )\n"
#endif
            ;
}

// One run stepped on the device one beta step at a time, writing the population dumps after
// every step (the reference writes them after the energies, before the averages).
int run_with_dumps(const pa_params& p, const char* dataDir, int prec, bool store_e, bool store_l,
                   std::vector<pa_row>& rows, int* nrows, uint64_t* is_end, int32_t* hist) {
  pa_run* run = nullptr;
  int rc = pa_run_create(&p, &run);
  if (rc) return rc;
  const int L = p.L, N = L * L, Nh = N / 2, Ld2 = L / 2;
  std::vector<int32_t> IE;
  std::vector<uint64_t> mask;
  std::vector<int8_t> pop;
  int last = 0;
  for (;;) {
    if ((rc = pa_run_step(run, 1))) break;
    int done = 0, total = 0;
    if ((rc = pa_run_steps_done(run, &done, &total))) break;
    if (done == last) break;  // schedule finished
    last = done;
    char path[700];
    if (store_e) {
      size_t bytes = 0;
      if ((rc = pa_run_array(run, 0, nullptr, &bytes))) break;
      IE.resize(bytes / 4);
      if ((rc = pa_run_array(run, 0, IE.data(), &bytes))) break;
      snprintf(path, sizeof path, "%sPA_energies_%d.dat", dataDir, done);
      std::ofstream f(path);
      f.precision(prec);
      if (p.coding == PA_SSC) {
        for (size_t j = 0; j < IE.size(); ++j) f << IE[j] << " ";
      } else {  // active lanes, group-major (PAisingMSC.cu:754-756)
        if ((rc = pa_run_array(run, 4, nullptr, &bytes))) break;
        mask.resize(bytes / 8);
        if ((rc = pa_run_array(run, 4, mask.data(), &bytes))) break;
        for (size_t g = 0; g < mask.size(); ++g)
          for (int l = 0; l < p.msbits; ++l)
            if ((mask[g] >> l) & 1) f << IE[g * p.msbits + l] << " ";
      }
    }
    if (store_l) {
      size_t bytes = 0;
      if ((rc = pa_run_population(run, nullptr, &bytes))) break;
      pop.resize(bytes);
      if ((rc = pa_run_population(run, pop.data(), &bytes))) break;
      pa_row cur[1];
      int nr = 0;
      std::vector<pa_row> tmp(done + 1);
      if ((rc = pa_run_results(run, tmp.data(), done + 1, &nr))) break;
      cur[0] = tmp[done];
      snprintf(path, sizeof path, "%sPA_lattice_%d.dat", dataDir, done);
      std::ofstream f(path, std::ofstream::out | std::ofstream::app);
      f.precision(prec);
      f << "i=" << done << "; beta=" << cur[0].B << "; R=" << cur[0].R << ";" << std::endl;
      const int R = cur[0].R;
      int ABflag = 0;
      for (int j = 0; j < R; ++j) {
        const int8_t* gA = &pop[(size_t)j * N];
        const int8_t* gB = gA + Nh;
        f << "P={";
        int k = 0;
        while (k < Nh) {
          if ((k % Ld2) == 0) { f << "{"; ABflag = 1 - ABflag; }
          if (ABflag) f << (int)gA[k] << "," << (int)gB[k];
          else f << (int)gB[k] << "," << (int)gA[k];
          k++;
          if ((k % Ld2) == 0) { f << "}"; if (k < Nh) f << ","; }
          else f << ",";
        }
        f << "};" << std::endl;
      }
      f << std::endl;
    }
  }
  if (!rc) rc = pa_run_results(run, rows.data(), (int)rows.size(), nrows);
  if (!rc) rc = pa_run_final_sequence(run, is_end);
  if (!rc && hist) rc = pa_run_histogram(run, hist, N + 1);
  pa_run_destroy(run);
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  const bool msc = PA_CLI_CODING == 1;
  const char* prog = msc ? "PAisingMSC" : "PAisingSSC";
  char dataDir[512];
  int optdir = 0;
  long long seed_opt = (long long)time(NULL);
  unsigned int EQsweeps = 100;
  double Bfin = 1.0, dBinit = 0.005;
  int Rinit = 20000, runs = 1, OutputPrecision = 11, L = 64, bits = 32, uconv = 0, adaptive = 0, mhr = 0, store_e = 0, store_l = 0;
  int optc, opti;
  double optf;
  while ((optc = get_opt(argc, argv, msc ? "R:t:d:f:M:s:P:o:L:w:u:aHe?" : "R:t:d:f:M:s:P:o:L:u:aHel?")) != -1) {
    switch (optc) {
      case 'R': opti = atoi(opt_arg); if (opti) Rinit = opti; break;
      case 't': opti = atoi(opt_arg); EQsweeps = opti; break;
      case 'd': optf = atof(opt_arg); if (optf > 0.0) dBinit = optf; break;
      case 'f': optf = atof(opt_arg); if (optf > 0.0) Bfin = optf; break;
      case 'M': opti = atoi(opt_arg); if (opti) runs = opti; break;
      case 's': opti = atoi(opt_arg); if (opti) seed_opt = opti; break;
      case 'P': opti = atoi(opt_arg); if (opti) OutputPrecision = opti; break;
      case 'o':
        if (opt_arg[strlen(opt_arg) - 1] == '/') snprintf(dataDir, sizeof dataDir, "%s", opt_arg);
        else snprintf(dataDir, sizeof dataDir, "%s/", opt_arg);
        optdir = 1;
        break;
      case 'L': opti = atoi(opt_arg); if (opti) L = opti; break;
      case 'w': opti = atoi(opt_arg); if (opti) bits = opti; break;
      case 'u': uconv = atoi(opt_arg); break;
      case 'a': adaptive = 1; break;
      case 'H': mhr = 1; break;
      case 'e': store_e = 1; break;
      case 'l': store_l = 1; break;
      case '?': usage(prog); return 1;
    }
  }
  if (opt_ind < argc) {
    for (opti = opt_ind; opti < argc; opti++) fprintf(stderr, "Non-option argument %s\n", argv[opti]);
    return 1;
  }
  // SSC keeps the seed in an int (PAisingSSC.cu:384), MSC in an unsigned long long
  // (PAisingMSC.cu:549); both reach curand_init as unsigned long long.
  const int seed_int = (int)seed_opt;
  const unsigned long long seed = msc ? (unsigned long long)seed_opt : (unsigned long long)(long long)seed_int;
  if (!optdir) {
    if (adaptive)  // PAisingSSC.cu:406-408: no dB in the name of an adaptive run
      snprintf(dataDir, sizeof dataDir, "./data%s_L%d_R%d_EqSw%d/", msc ? "MSC" : "SSC", L, Rinit,
               EQsweeps);
    else
      snprintf(dataDir, sizeof dataDir, "./data%s_L%d_R%d_EqSw%d_dB%f/", msc ? "MSC" : "SSC", L,
               Rinit, EQsweeps, dBinit);
  }
  mkdir(dataDir, 0777);

  pa_params p;
  memset(&p, 0, sizeof p);
  p.L = L;
  p.coding = msc ? PA_MSC : PA_SSC;
  p.msbits = bits;
  p.Rinit = Rinit;
  p.eqsweeps = EQsweeps;
  p.dB = dBinit;
  p.Binit = 0.0;
  p.Bfin = Bfin;
  p.seed = seed;
  p.uniform = uconv ? PA_UNIFORM_CUDA100 : PA_UNIFORM_CUDA101;

  if (msc) std::cout << "RNG initial seed: " << seed << "\n";
  else std::cout << "RNG initial seed: " << seed_int << "\n";
  const size_t fp = pa_run_footprint(&p);
  const int N = L * L;
  if (fp == 0) {
    fprintf(stderr, "%s\n", pa_last_error());
    return 1;
  }
  std::cout << "Memory use of one replica: " << (double)fp / Rinit / 1024.0 << " kB \n";
  std::cout << "Memory use of the entire population of " << Rinit
            << " replicas: " << (double)fp / 1024.0 / 1024.0 << " MB \n";
  fflush(stdout);

  // The reference creates its device context (cudaDeviceSetCacheConfig, cudaEventCreate)
  // before its timer starts (PAisingSSC.cu:439-445); initialise the device the same way.
  {
    char name[128];
    int cus = 0;
    size_t hbm = 0;
    if (pa_device_info(name, sizeof name, &cus, &hbm) != PA_OK) {
      fprintf(stderr, "%s\n", pa_last_error());
      return 1;
    }
  }
  const auto t0 = std::chrono::steady_clock::now();
  double totPop = 0.0;
  unsigned long long IS = 0;
  std::vector<pa_row> rows(10000);
  std::vector<double> sched;  // adaptive: run 0's betas, replayed by later runs (SSC:662)
  p.adaptive = adaptive;
  p.mhr = mhr;
  std::vector<int32_t> hist(mhr ? N + 1 : 0);
  // MHR overwrites BF[0] (set once before the runs, SSC:435) and the next run prints and
  // uses the overwritten value (SSC:678-690, 711)
  double BF0 = -N * std::log(2.0);
  for (int r = 0; r < runs; ++r) {
    p.initial_sequence = IS;
    if (adaptive && r > 0) {
      p.adaptive = 0;
      p.schedule = sched.data();
      p.schedule_len = (int)sched.size();
    }
    int nrows = 0;
    uint64_t is_end = 0;
    int rc;
    if (store_e || (store_l && r == 0))
      rc = run_with_dumps(p, dataDir, OutputPrecision, store_e, store_l && r == 0, rows, &nrows,
                          &is_end, mhr ? hist.data() : nullptr);
    else if (mhr)
      rc = pa_run_all_hist(&p, rows.data(), (int)rows.size(), &nrows, &is_end, hist.data(), N + 1);
    else
      rc = pa_run_all(&p, rows.data(), (int)rows.size(), &nrows, &is_end);
    if (rc == PA_ERR_NBMAX) {
      if (adaptive)
        fprintf(stderr,
                "Error: number of temperature steps exceeds nBmax=%d.\n Please consider increasing the "
                "population size or decreasing the value of MinOverlap or increasing the value of nBmax.\n",
                10000);
      else
        fprintf(stderr,
                "Error: number of temperature steps exceeds nBmax=%d.\n Please consider increasing the "
                "inverse temperature step or increasing the value of nBmax.\n",
                10000);
      return 1;
    }
    if (rc) {
      fprintf(stderr, "%s\n", pa_last_error());
      return rc;
    }
    IS = is_end;
    if (adaptive && r == 0)
      for (int i = 0; i < nrows; ++i) sched.push_back(rows[i].B);
    for (int i = 1; i < nrows; ++i) totPop += rows[i].R;
    if (mhr) rows[0].BF = BF0 / N;
    char str[700];
    snprintf(str, sizeof str, "%sPA_results_run_%d.dat", dataDir, r);
    std::ofstream results(str);
    results.precision(OutputPrecision);
    for (int i = 0; i < nrows; ++i) {
      const pa_row& w = rows[i];
      results << w.B << " " << w.E << " " << w.C << " " << w.M << " " << w.M2 << " " << w.M4 << " "
              << w.BF << " " << w.S << " " << w.R << " " << w.lnQ << "\n";
    }
    if (mhr) {
      // BF_i = BF_0 + sum of -lnQ in the reference's order (SSC:627-628), R_i, B_i
      std::vector<double> B(nrows), BF(nrows);
      std::vector<int32_t> R(nrows);
      double sumlnQ = 0.0;
      for (int i = 0; i < nrows; ++i) {
        B[i] = rows[i].B;
        R[i] = rows[i].R;
        if (i == 0) { BF[i] = BF0; continue; }
        sumlnQ -= rows[i].lnQ;
        BF[i] = -N * std::log(2.0) + sumlnQ;
      }
      const int niter = 1;  // MHR_Niter, SSC:53
      std::vector<double> out((size_t)nrows * (1 + 3 * niter));
      if (pa_mhr_analysis(N, nrows, B.data(), BF.data(), R.data(), hist.data(), niter, out.data())) {
        fprintf(stderr, "%s\n", pa_last_error());
        return 1;
      }
      BF0 = BF[0];
      snprintf(str, sizeof str, "%sPA_MHR_results_run_%d.dat", dataDir, r);
      std::ofstream mres(str);
      mres.precision(OutputPrecision);
      for (int i = 0; i < nrows; ++i) {
        const double* o = &out[(size_t)i * (1 + 3 * niter)];
        mres << o[0] << " ";
        for (int l = 0; l < niter; ++l) mres << o[1 + 3 * l] << " " << o[2 + 3 * l] << " " << o[3 + 3 * l] << " ";
        mres << "\n";
      }
    }
  }
  const double Etime =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "Elapsed time: " << std::setprecision(8) << Etime / 1000 << " s\n";
  std::cout << "Time per spin-flip: " << std::setprecision(8) << Etime * 1e6 / EQsweeps / N / totPop
            << " ns\n";
  return 0;
}
