/*
 * tb8gen_main.c -- C host driver of the MI355X king-slice engine.
 *
 * Keeps the reference generator's command line and stage order (main(),
 * src/tb8gen.c:1245-1453): parse the material, generate every smaller
 * material its captures reach (the reference probes Syzygy files there,
 * src/probe.c:1522-1684; here the engine's own tables are resident on the
 * device), resolve capture exits, illegal/mate, then the ply loop with the
 * reference's more_* flags, printing the same count lines.
 *
 * Usage: tb8gen [-f plain|batched|graph] [-d device] [-v] MATERIAL
 *   -v   also print the count lines of the smaller materials
 */
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "tbg.h"

#define MAXTAB 256

static tbg_ctx *g_ctx[MAXTAB];
static char g_name[MAXTAB][16];
static int g_ntab;

static const char *kOrder = "KQRBN";

static void die(const char *what, int err)
{
  fprintf(stderr, "%s: %s\n", what, tbg_strerror(err));
  exit(EXIT_FAILURE);
}

static void print_line(const char *line, void *user)
{
  if (user) printf("%s\n", line);
}

/* the side with more pieces, then the stronger piece string, is white */
static int side_cmp(const char *a, size_t la, const char *b, size_t lb)
{
  if (la != lb) return la > lb ? 1 : -1;
  for (size_t i = 0; i < la; i++)
    if (a[i] != b[i]) return strchr(kOrder, a[i]) < strchr(kOrder, b[i]) ? 1 : -1;
  return 0;
}

/* material string with one piece (index `skip` of the letters, -1 none),
 * canonical colour orientation */
static void sub_name(const char *m, int skip, char *out)
{
  char w[16], b[16];
  int nw = 0, nb = 0, side = 0, i = 0;
  for (const char *s = m; *s; s++) {
    if (*s == 'v') { side = 1; continue; }
    if (i++ == skip) continue;
    if (side) b[nb++] = *s; else w[nw++] = *s;
  }
  /* order each side K Q R B N (create_material_string, probe.c:1480) */
  for (int x = 0; x < nw; x++)
    for (int y = x + 1; y < nw; y++)
      if (strchr(kOrder, w[y]) < strchr(kOrder, w[x])) { char t = w[x]; w[x] = w[y]; w[y] = t; }
  for (int x = 0; x < nb; x++)
    for (int y = x + 1; y < nb; y++)
      if (strchr(kOrder, b[y]) < strchr(kOrder, b[x])) { char t = b[x]; b[x] = b[y]; b[y] = t; }
  w[nw] = b[nb] = 0;
  if (side_cmp(w, nw, b, nb) >= 0) snprintf(out, 16, "%sv%s", w, b);
  else snprintf(out, 16, "%sv%s", b, w);
}

static int find(const char *name)
{
  for (int i = 0; i < g_ntab; i++)
    if (!strcmp(g_name[i], name)) return i;
  return -1;
}

static int count_pieces(const char *m)
{
  int n = 0;
  for (const char *s = m; *s; s++) n += *s != 'v';
  return n;
}

/* generate `name` after every smaller material it needs */
static int generate(const char *name, int form, int verbose, int top)
{
  int h = find(name);
  if (h >= 0) return h;
  int letters = count_pieces(name);
  for (int i = 0; i < letters; i++) {
    /* skip the kings */
    int idx = 0, side = 0, is_king = 0;
    for (const char *s = name; *s; s++) {
      if (*s == 'v') { side = 1; continue; }
      if (idx++ == i) { is_king = *s == 'K'; break; }
    }
    (void)side;
    if (is_king) continue;
    char sub[16];
    sub_name(name, i, sub);
    if (count_pieces(sub) <= 2) continue;
    generate(sub, form, verbose, 0);
  }
  if (g_ntab >= MAXTAB) { fprintf(stderr, "too many tables\n"); exit(EXIT_FAILURE); }
  tbg_ctx *c;
  int r = tbg_create(&c, name);
  if (r) die(name, r);
  r = tbg_resolve_captures(c, g_ctx, g_ntab);
  if (r) die(name, r);
  tbg_counts cnt;
  r = tbg_generate(c, form, print_line, (top || verbose) ? (void *)1 : NULL, &cnt);
  if (r) die(name, r);
  g_ctx[g_ntab] = c;
  snprintf(g_name[g_ntab], 16, "%s", name);
  return g_ntab++;
}

int main(int argc, char **argv)
{
  int opt, form = TBG_FORM_BATCHED, device = 0, verbose = 0;
  while ((opt = getopt(argc, argv, "f:d:t:v")) != -1) {
    switch (opt) {
    case 'f':
      form = !strcmp(optarg, "plain") ? TBG_FORM_PLAIN
           : !strcmp(optarg, "graph") ? TBG_FORM_GRAPH
                                      : TBG_FORM_BATCHED;
      break;
    case 'd': device = atoi(optarg); break;
    case 't': break; /* accepted for compatibility: the device replaces the thread pool */
    case 'v': verbose = 1; break;
    default: return 2;
    }
  }
  if (optind >= argc) {
    printf("No tablebase specified.\n");
    return EXIT_FAILURE;
  }
  int r = tbg_init(device);
  if (r) die("tbg_init", r);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  generate(argv[optind], form, verbose, 1);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  fprintf(stderr, "generated %d tables in %.3f s\n", g_ntab,
          (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec));
  for (int i = 0; i < g_ntab; i++) tbg_destroy(g_ctx[i]);
  return 0;
}
