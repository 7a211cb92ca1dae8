// CPU oracle for the counting-sort chunker + octree indexer path.
//
// TEST INFRASTRUCTURE ONLY.  This library is a single-threaded restatement
// of the reference algorithm (PotreeConverter 2.x) used as the parity
// checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg.  The product (potreeconverter_amd) never links, loads or calls it.
//
// Every function cites the reference file:line it restates.  Schedules that
// are nondeterministic in the reference (chunk append order, indexer thread
// completion order, hierarchy batch directory order) are replaced by the
// canonical order: input order inside chunks, chunks in ascending ID,
// per-chunk post-order completions, phase-2 tasks in pre-order, root last,
// hierarchy batches in ascending name.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// A3: per-cell point counts (chunker_countsort_laszip.cpp:131-361).
// las: n records of bpp bytes (X,Y,Z int32 at offset 0).  Returns 0, or 123
// when a point lies outside the cubic box (:236-251).
int orc_count_cells(const uint8_t* las, int64_t n, int32_t bpp, const double las_scale[3],
                    const double las_offset[3], const double pos_scale[3], const double pos_offset[3],
                    const double min[3], const double max[3], int32_t grid_size, int32_t* grid);

// A4: chunk list + LUT (chunker_countsort_laszip.cpp:1074-1204).  Writes up to
// max_chunks chunk descriptors (level, x, y, z, count) in the reference's
// push order and the LUT (grid_size^3 int32, -1 = no chunk).  Returns the
// chunk count (may exceed max_chunks; then only the count is valid).
int64_t orc_create_lut(const int32_t* grid, int32_t grid_size, int64_t max_points_per_chunk,
                       int64_t* chunk_desc /* 5 per chunk */, int64_t max_chunks, int32_t* lut);

// Whole conversion of one LAS file: doChunking + doIndexing with the given
// method ("poisson" | "poisson_average" | "random"), canonical order.  Writes
// outdir/{octree.bin, hierarchy.bin, metadata.json}; with keep_chunks also
// outdir/chunks/{<id>.bin, metadata.json}.  chunk_only stops after the
// chunker.  timings (may be null) receives
// [count, lut+distribute, chunking total, indexing] seconds.
int orc_convert(const char* las_path, const char* outdir, const char* method, int32_t keep_chunks,
                int32_t chunk_only, double* timings);

// orc_convert over several LAS files with the --attributes filter
// (computeOutputAttributes(sources, requestedAttributes),
// PotreeConverter.h:195-315) and the chunker's box (min / max from
// computeStats, main.cpp:215-284; both null: computed from the headers).
int orc_convert_sources(const char* const* las_paths, int32_t num_sources, const double* min, const double* max,
                        const char* const* attributes, int32_t num_attributes, const char* outdir,
                        const char* method, int32_t keep_chunks, int32_t chunk_only, double* timings);

// doIndexing alone on an existing outdir/chunks directory written by the
// chunker (chunks are read in ascending ID order).
int orc_index(const char* outdir, const char* method, const char* name, double* timings);
// orc_convert with the node encoding of the reference's --encoding
// ("BROTLI" compresses the nodes; any other name writes the records
// unchanged and is echoed in metadata.json)
int orc_convert_encoded(const char* las_path, const char* outdir, const char* method, const char* encoding,
                        int32_t keep_chunks, int32_t chunk_only, double* timings);

// Poisson node sampling on one synthetic node (sampler_poisson.h:14-304 /
// sampler_poisson_average.h:14-492) for unit tests: children point buffers
// are concatenated (bpp bytes each, counts per child, -1 = no child,
// -2 = child is an inner node).  Outputs accept flags in child-then-index
// order and the (possibly recoloured) child records.
int orc_sample_node(const char* method, const uint8_t* records, const int64_t child_counts[8], int32_t bpp,
                    int32_t rgb_offset, const double node_min[3], const double node_max[3], int32_t level,
                    double base_spacing, const double pos_scale[3], const double pos_offset[3],
                    uint8_t* accepted_flags, uint8_t* out_records);

// laszip point fields of n records (32-byte entries, see ref_point_decode.cpp)
int orc_decode_points(int32_t fmt, int32_t record_length, const uint8_t* records, int64_t n, uint8_t* out);
const char* orc_last_error(void);

// Seed of the method "random" (SamplerRandom, sampler_random.h): every leaf
// shuffle uses std::default_random_engine(seed), the value the reference
// takes from system_clock::now() (sampler_random.h:102-104).
int orc_set_random_seed(uint32_t seed);
/* Forces the chunk grid size and chunk cap of the next conversions (0: the
 * reference's rule from the point count). */
int orc_set_chunking(int64_t grid, int64_t cap);

#ifdef __cplusplus
}
#endif
