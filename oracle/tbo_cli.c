/* tbo_cli -- command-line front end of the CPU oracle (test infrastructure).
 * Usage: tbo_cli [-t threads] [-k] MATERIAL
 *        tbo_cli [-t threads] -s SLICE -S SEED [-p MAXPLIES] MATERIAL
 * Prints the reference's count lines (tb8gen.c:764-767, 1024, 1082). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#include <time.h>
#include "tbo.h"

int main(int argc, char **argv)
{
  int opt, threads = 1, slice = -1, maxp = 0;
  unsigned long long seed = 0;
  while ((opt = getopt(argc, argv, "t:ks:S:p:")) != -1) {
    switch (opt) {
    case 't': threads = atoi(optarg); break;
    case 'k': tbo_set_ref_king_capture_index(1); break;
    case 's': slice = atoi(optarg); break;
    case 'S': seed = strtoull(optarg, NULL, 0); break;
    case 'p': maxp = atoi(optarg); break;
    default: return 2;
    }
  }
  if (optind >= argc) { fprintf(stderr, "No tablebase specified.\n"); return 1; }
  tbo_set_threads(threads);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int h = slice >= 0 ? tbo_synth_slice(argv[optind], slice, seed, "-", maxp)
                     : tbo_generate(argv[optind], "-");
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (h < 0) { fprintf(stderr, "generation failed (%d)\n", h); return 1; }
  fprintf(stderr, "time %.3f s, plies %d\n",
          (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec), tbo_num_plies(h));
  return 0;
}
