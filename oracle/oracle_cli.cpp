// Command-line front end of the CPU oracle (test infrastructure only).
//   pc_oracle convert <file.las> <outdir> <poisson|poisson_average> [keep_chunks]
//   pc_oracle index <outdir> <poisson|poisson_average> <name>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>

#include "oracle.h"

int main(int argc, char** argv) {
  if (argc >= 5 && !strcmp(argv[1], "convert")) {
    double t[4] = {0, 0, 0, 0};
    int rc = orc_convert(argv[2], argv[3], argv[4], argc > 5 ? atoi(argv[5]) : 0, 0, t);
    if (rc) fprintf(stderr, "error %d: %s\n", rc, orc_last_error());
    printf("count %.3f distribute %.3f chunking %.3f indexing %.3f\n", t[0], t[1], t[2], t[3]);
    return rc;
  }
  if (argc >= 5 && !strcmp(argv[1], "index")) {
    double t[1] = {0};
    int rc = orc_index(argv[2], argv[3], argv[4], t);
    if (rc) fprintf(stderr, "error %d: %s\n", rc, orc_last_error());
    printf("indexing %.3f\n", t[0]);
    return rc;
  }
  fprintf(stderr, "usage: pc_oracle convert <las> <outdir> <method> [keep] | index <outdir> <method> <name>\n");
  return 2;
}
