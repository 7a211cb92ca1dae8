// Test-only driver that runs the reference indexer (indexer::doIndexing with
// SamplerPoisson / SamplerPoissonAverage / SamplerRandom) on an existing <outdir>/chunks
// directory.  It is compiled against the reference sources in place by
// oracle/build_ref.sh and lives only in oracle/_ref/ (never shipped).
//   ref_indexer <outdir> <poisson|poisson_average|random> <name> <pointsTotal> [DEFAULT|BROTLI]
#include <stdlib.h>

#include <iostream>
#include <string>

#include <chrono>

#include "indexer.h"
#include "logger.h"
#include "sampler_poisson.h"
#include "sampler_poisson_average.h"
#include "sampler_random.h"

// SamplerRandom seeds every leaf shuffle with system_clock::now()
// (sampler_random.h:102-104).  This driver fixes the clock at REF_CLOCK_NS
// nanoseconds since the epoch (environment, default 0), which makes "random"
// the seeded variant that the oracle and the GPU path implement.  The
// reference reads the clock otherwise only for progress timing.
namespace std {
namespace chrono {
inline namespace _V2 {
system_clock::time_point system_clock::now() noexcept {
  const char* e = getenv("REF_CLOCK_NS");
  return time_point(duration(e ? strtoll(e, nullptr, 10) : 0));
}
}  // namespace _V2
}  // namespace chrono
}  // namespace std

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
  } else if (method == "random") {
    SamplerRandom sampler;
    indexer::doIndexing(outdir, state, options, sampler);
  } else {
    std::cerr << "unknown method " << method << std::endl;
    return 2;
  }
  return 0;
}
