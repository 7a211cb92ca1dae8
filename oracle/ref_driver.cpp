// Test-only driver that runs the reference indexer (indexer::doIndexing with
// SamplerPoisson / SamplerPoissonAverage / SamplerRandom) on an existing <outdir>/chunks
// directory.  It is compiled against the reference sources in place by
// oracle/build_ref.sh and lives only in oracle/_ref/ (never shipped).
//   ref_indexer <outdir> <poisson|poisson_average|random> <name> <pointsTotal> [DEFAULT|BROTLI]
// Built a second time with REF_GPU_SAMPLERS as ref_indexer_gpu: the same
// reference indexer with the GPU samplers plugged in through the Sampler
// interface (libpotree_gpu.so, pcg_sample_subtree).
#include <stdlib.h>

#include <iostream>
#include <string>

#include <chrono>

#include "indexer.h"
#include "logger.h"
#include "sampler_poisson.h"
#include "sampler_poisson_average.h"
#include "sampler_random.h"
#ifdef REF_GPU_SAMPLERS
// the GPU samplers of the integration guide (PotreeGpuSampler.h, taken
// from INTEGRATION.md by oracle/build_ref.sh) in place of the reference's
#include "PotreeGpuSampler.h"
#define SamplerPoisson SamplerPoissonGPU
#define SamplerPoissonAverage SamplerPoissonAverageGPU
#define SamplerRandom SamplerRandomGPU
#endif

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
#ifdef REF_GPU_SAMPLERS
  // the leaf shuffles of the GPU SamplerRandom take the library's seed
  const char* clk = getenv("REF_CLOCK_NS");
  if (pcg_set_option("random_seed", (int64_t)(uint32_t)(clk ? strtoll(clk, nullptr, 10) : 0)) != 0) return 3;
#endif
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
