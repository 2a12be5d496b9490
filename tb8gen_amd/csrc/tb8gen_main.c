/*
 * tb8gen_main.c -- C host driver of the MI355X king-slice engine.
 *
 * Keeps the reference generator's command line and stage order (main(),
 * src/tb8gen.c:1245-1453): parse the material, generate every smaller
 * material its captures reach with the library's scheduler (the reference
 * probes Syzygy files there, src/probe.c:1522-1684; here the engine's own
 * tables are resident on the device), resolve capture exits, illegal/mate,
 * then the ply loop with the reference's more_* flags, printing the same
 * count lines.
 *
 * Usage: tb8gen [-f plain|batched|graph] [-d device] [-v] [-w DIR] MATERIAL
 *   -v   also print the count lines of the smaller materials
 *   -w   write the finished tables (the material and every smaller one)
 *        as Syzygy WDL files DIR/<name>.rtbw, the files the reference's
 *        probe_wdl reads (the merge() stub of src/tb8gen.c:1448-1452)
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "tbg.h"

static void die(const char *what, int err)
{
  fprintf(stderr, "%s: %s\n", what, tbg_strerror(err));
  exit(EXIT_FAILURE);
}

static void print_line(const char *line, void *user)
{
  (void)user;
  printf("%s\n", line);
}

int main(int argc, char **argv)
{
  int opt, form = TBG_FORM_BATCHED, device = 0, verbose = 0, threads = 1;
  const char *wdl_dir = NULL;
  while ((opt = getopt(argc, argv, "f:d:t:vw:")) != -1) {
    switch (opt) {
    case 'f':
      form = !strcmp(optarg, "plain") ? TBG_FORM_PLAIN
           : !strcmp(optarg, "graph") ? TBG_FORM_GRAPH
                                      : TBG_FORM_BATCHED;
      break;
    case 'd': device = atoi(optarg); break;
    case 't': threads = atoi(optarg); break; /* printed only: the device replaces the pool */
    case 'v': verbose = 1; break;
    case 'w': wdl_dir = optarg; break;
    default: return 2;
    }
  }
  if (optind >= argc) {
    printf("No tablebase specified.\n");
    return EXIT_FAILURE;
  }
  int r = tbg_init(device);
  if (r) die("tbg_init", r);
  /* the reference's first line (tb8gen.c:1322-1324) */
  printf("number of threads = %d\n", threads > 1 ? threads : 1);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  /* the material and every smaller one its captures reach, smallest first,
     one batch per piece count (the library's scheduler) */
  tbg_registry *reg;
  r = tbg_registry_create(&reg, form);
  if (r) die("registry", r);
  tbg_ctx *ctx;
  r = tbg_registry_get(reg, argv[optind], &ctx);
  if (r) die(argv[optind], r);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  int n = tbg_registry_count(reg);
  for (int i = 0; verbose && i < n; i++) {
    const char *name;
    tbg_ctx *c;
    if (tbg_registry_entry(reg, i, &name, &c) == 0 && c != ctx) {
      printf("# %s\n", name);
      tbg_emit_lines(c, print_line, NULL);
    }
  }
  r = tbg_emit_lines(ctx, print_line, NULL);
  if (r) die(argv[optind], r);
  fprintf(stderr, "generated %d tables in %.3f s\n", n,
          (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec));
  for (int i = 0; wdl_dir && i < n; i++) {
    const char *name;
    tbg_ctx *c;
    char path[4096];
    if (tbg_registry_entry(reg, i, &name, &c)) continue;
    r = tbg_write_rtbw(c, wdl_dir, path, sizeof path);
    if (r) die(name, r);
    fprintf(stderr, "wrote %s\n", path);
  }
  tbg_registry_destroy(reg);
  return 0;
}
