// Test-only driver that runs the reference indexer (indexer::doIndexing with
// SamplerPoisson / SamplerPoissonAverage) on an existing <outdir>/chunks
// directory.  It is compiled against the reference sources in place by
// oracle/build_ref.sh and lives only in oracle/_ref/ (never shipped).
//   ref_indexer <outdir> <poisson|poisson_average> <name> <pointsTotal> [DEFAULT|BROTLI]
#include <stdlib.h>

#include <iostream>
#include <string>

#include "indexer.h"
#include "logger.h"
#include "sampler_poisson.h"
#include "sampler_poisson_average.h"

int main(int argc, char** argv) {
  if (argc < 5) {
    std::cerr << "usage: ref_indexer <outdir> <method> <name> <pointsTotal>" << std::endl;
    return 2;
  }
  std::string outdir = argv[1], method = argv[2];
  State state;
  state.pointsTotal = atoll(argv[4]);
  Options options;
  options.name = argv[3];
  options.method = method;
  options.encoding = argc > 5 ? argv[5] : "DEFAULT";
  options.keepChunks = true;
  logger::addOutputFile(outdir + "/log.txt");
  if (method == "poisson") {
    SamplerPoisson sampler;
    indexer::doIndexing(outdir, state, options, sampler);
  } else if (method == "poisson_average") {
    SamplerPoissonAverage sampler;
    indexer::doIndexing(outdir, state, options, sampler);
  } else {
    std::cerr << "unknown method " << method << std::endl;
    return 2;
  }
  return 0;
}
