// Test-only cross-check: words of rocRAND's Philox4x32-10 engine (MIT licensed,
// /opt/rocm/include/rocrand/rocrand_philox4x32_10.h) for given (seed, subsequence, offset),
// evaluated on the host.  Usage: rocrand_stream_words seed subsequence offset count
#include <cstdio>
#include <cstdlib>
#include <rocrand/rocrand_kernel.h>

int main(int argc, char** argv) {
  if (argc != 5) return 2;
  unsigned long long seed = strtoull(argv[1], 0, 0), sub = strtoull(argv[2], 0, 0),
                     off = strtoull(argv[3], 0, 0);
  int n = atoi(argv[4]);
  rocrand_state_philox4x32_10 st;
  rocrand_init(seed, sub, off, &st);
  for (int i = 0; i < n; ++i) printf("%u\n", rocrand(&st));
  return 0;
}
