/* A plain C program against include/potree_gpu.h: generates a seeded LAS
 * tile with the library's generator, converts it (both phases, the
 * reference converter's main.cpp:353 and :380-390), then converts it again
 * through the separate phase entry points, and prints the timings.
 *   convert <work dir> <points> <seed> <poisson|poisson_average|random>
 * Exit status 0 on success; errors print pcg_last_error(). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "potree_gpu.h"

static int fail(const char* what) {
  fprintf(stderr, "%s: %s\n", what, pcg_last_error());
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 5) {
    fprintf(stderr, "usage: convert <work dir> <points> <seed> <method>\n");
    return 2;
  }
  const char* dir = argv[1];
  const int64_t n = atoll(argv[2]);
  const uint64_t seed = strtoull(argv[3], NULL, 10);
  const char* method = argv[4];
  char las[4096], out[4096], split[4096];
  snprintf(las, sizeof las, "%s/in.las", dir);
  snprintf(out, sizeof out, "%s/out", dir);
  snprintf(split, sizeof split, "%s/split", dir);
  if (pcg_init() != 0) return fail("pcg_init");
  unsigned char params[256];
  if (pcg_lasgen_preset(0, n, seed, params) != 0) return fail("pcg_lasgen_preset");
  if (pcg_lasgen_write(params, las) != 0) return fail("pcg_lasgen_write");
  if (strcmp(method, "random") == 0 && pcg_set_option("random_seed", 12345) != 0) return fail("pcg_set_option");
  double t[8];
  if (pcg_convert(las, out, method, 0, t) != 0) return fail("pcg_convert");
  printf("convert: chunking %.4f s, indexing %.4f s\n", t[2], t[3]);
  if (pcg_do_chunking(las, split, t) != 0) return fail("pcg_do_chunking");
  if (pcg_do_indexing(split, method, "in", n, t) != 0) return fail("pcg_do_indexing");
  printf("split phases: indexing %.4f s\n", t[3]);
  /* an error path: a missing file is reported, not fatal */
  if (pcg_convert("/nonexistent.las", out, method, 0, NULL) == 0) return fail("missing input accepted");
  printf("missing input: %s\n", pcg_last_error());
  return pcg_release();
}
