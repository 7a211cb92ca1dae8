/*
 * C ABI of the MI355X (gfx950) counting-sort chunker and octree indexer.
 *
 * Plain pointers, sizes and status codes only.  Every entry point returns 0
 * on success; 123 mirrors the reference converter's fatal exit code (point
 * outside the bounding box, failed chunk lookup, bad hierarchy); any other
 * nonzero value is an error whose text pcg_last_error() returns.  Device
 * pointers are HIP device allocations on device 0; `stream` arguments are
 * hipStream_t values passed as void* (NULL = the library's own stream).
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the PotreeConverter 2.x checkout).
 */
#ifndef POTREE_GPU_H
#define POTREE_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Selects HIP device 0 and creates the library stream.  Called implicitly by
 * every other entry point. */
int pcg_init(void);

/* Text of the last error (thread-unsafe, like the reference's logger). */
const char* pcg_last_error(void);

/* The library keeps its device buffers (chunk records, sampling scratch,
 * octree image, ...) between calls so that a job reuses the previous job's
 * memory.  pcg_release frees them; the next call allocates again.  No
 * reference counterpart (the reference converter is a one-shot process). */
int pcg_release(void);

/* Library options.  "plain_sampling" (0/1): 1 launches the sampling kernels
 * once per octree node instead of once per level for all its nodes (same
 * results; the unbatched baseline the batched path is measured against).
 * "ring_min_nodes" (default 512): Poisson octree levels with at least this
 * many nodes decide acceptance with the accepted-shell ring kernel, the
 * others with the cell-table kernels (same results).  "sample_groups"
 * (0-8, default 0): the chunk subtrees are sampled in this many sets, each
 * on a stream of its own, before the nodes above the chunk roots (same
 * results; 1 = one level-synchronous pass over the whole tree; 0 = one set
 * per hardware queue besides the library stream's, GPU_MAX_HW_QUEUES - 1).
 * "random_seed" (-1 or 0 .. 2^32-1, default -1): seed of every leaf shuffle
 * of -m random (SamplerRandom, Converter/include/sampler_random.h:102-104,
 * where it is system_clock::now()); -1 takes the clock as the reference does.
 * "sample_sets" (0-64, default 0): the chunk subtrees are split into this
 * many sets (at least one per stream), the streams taking them in turn with
 * buffers sized for their largest set (same results; 0 = one set per stream
 * unless the job needs more sets to fit in device memory, as at 1B points).
 * "chunk_grid" (0, 128, 256, 512) and "chunk_cap" (0 or >= 1) force the
 * chunker's grid size and per-chunk point cap, which the reference derives
 * from the point count (chunker_countsort_laszip.cpp:1209-1220; 0 = that
 * rule): a small input then takes the paths of a large one.
 * "distribute_tail" (0/1, default 1): 1 places points with the tile
 * transcode (16-bit chunk ids, aligned 16-byte run stores) where the record
 * layout allows, 0 with the byte-wise scatter (same results).
 * "sample_order" (0/1, default 0): launch rounds of the node sampling by
 * octree level, deepest first (0), or by height, a node in the first round
 * after all its children (1) (same results).
 * "sample_graph" (0/1, default 0): 1 captures the sampling of the chunk
 * subtrees (every set's launch sequence on its stream, forked from and
 * joined to the library stream) in one graph per job and launches it (same
 * results).
 * Launch shapes of the node sampling (same results for every value):
 * "ring_blocks_per_cu" (1-3, default 3) persistent ring blocks per CU,
 * "copy_blocks" (default 8192) and "neighbor_blocks" (default 4096) grids of
 * the level-wide record copy and neighbour search, "avg_blocks" (default
 * 2048) workgroups of a Poisson-average colour-sum launch,
 * "bucket_min_nodes" (default 512) launches of at least this many nodes sort
 * each node's candidates by distance buckets, smaller ones by radix passes,
 * "accept_pipe_nodes" (default 1024, 0 = never) launches of at most this
 * many nodes, and every launch the ring decides, take the pipelined
 * acceptance (a workgroup per cell-path node), larger ones one wavefront
 * per node, "ring_max_cand" (default 2^30) nodes above this many candidates
 * take the cell-table path instead of the ring, "pipe_form" (0-2, default
 * 0; a test option) the pipelined acceptance by node size (0), every node in
 * its 32-bit-position form (1) or by one wavefront per node (2).
 * "set_policy" (0/1, default 0): chunk subtrees split into sets by their
 * points (0) or with the subtrees deeper than where most points end in sets
 * of their own (1) (same results).
 * The other options have no reference counterpart. */
int pcg_set_option(const char* key, int64_t value);

/* Current value of an option of pcg_set_option.  No reference counterpart. */
int pcg_get_option(const char* key, int64_t* value);

/* Autotuner of the runtime options above.  It converts the caller's
 * workload (records in HBM, as pcg_convert_device, no files written) once
 * per configuration and searches the options pcg_autotune_param_name(0 ..
 * pcg_autotune_params() - 1) names by coordinate descent: each option over
 * its candidate values with the others fixed, a value kept when the job is
 * at least 1 % faster, for at most `passes` rounds; a configuration's time
 * is the fastest of `repeats` conversions.  The best values stay set and are
 * written to chosen[] (pcg_autotune_params() entries); best_s and default_s
 * (nullable) receive the job's seconds with them and with the options as
 * they were.  Every option it tries gives the same output bytes.  No
 * reference counterpart (the reference has no launch parameters). */
int32_t pcg_autotune_params(void);
const char* pcg_autotune_param_name(int32_t i);
int pcg_autotune_device(const void* d_las, int64_t n, const uint8_t* las_header, const char* method,
                        int32_t passes, int32_t repeats, int64_t* chosen, double* best_s, double* default_s);

/* Phase 1 + phase 2 from a LAS file, the drop-in for the two calls at
 * Converter/src/main.cpp:353 (chunker_countsort_laszip::doChunking) and
 * Converter/src/main.cpp:380-390 (indexer::doIndexing with
 * SamplerPoisson / SamplerPoissonAverage / SamplerRandom).  method:
 * "poisson", "poisson_average" or "random".  Writes outdir/{octree.bin, hierarchy.bin,
 * metadata.json}; flags bit 0 keeps outdir/chunks/ (--keep-chunks).
 * timings (nullable, 8 doubles): chunking count, chunking distribute,
 * chunking total, indexing, octree build, sampling (seconds), nodes sorted
 * in libstdc++ order, nodes resolved sequentially. */
int pcg_convert(const char* las_path, const char* outdir, const char* method, int32_t flags, double* timings);

/* pcg_convert with the converter's --encoding option
 * (Converter/src/main.cpp:27,60,125, handed to indexer::doIndexing through
 * Options::encoding).  "BROTLI" writes every node as the reference's
 * compress() does (Converter/src/indexer.cpp:1110-1401: attribute arrays,
 * position and rgb as Morton codes, points in position-code order, brotli
 * quality 6); any other string writes the records unchanged
 * (indexer.cpp:1462) and is echoed in metadata.json (indexer.cpp:516).
 * NULL means "DEFAULT", the reference's default. */
int pcg_convert_encoded(const char* las_path, const char* outdir, const char* method, const char* encoding,
                        int32_t flags, double* timings);

/* Phase 1 only (chunker_countsort_laszip::doChunking,
 * Converter/include/chunker_countsort_laszip.h:19): writes outdir/chunks/
 * (<id>.bin + metadata.json) exactly as the reference chunker's contract. */
int pcg_do_chunking(const char* las_path, const char* outdir, double* timings);

/* doChunking with the reference's arguments
 * (chunker_countsort_laszip::doChunking(sources, targetDir, min, max, state,
 * outputAttributes, monitor), Converter/include/chunker_countsort_laszip.h:19,
 * called at Converter/src/main.cpp:353): num_sources LAS files; min / max
 * (both or neither, 3 doubles each) the box computeStats gives the chunker
 * (main.cpp:215-284; NULL: computed from the headers the same way);
 * attributes (num_attributes names, NULL / 0 for all) the --attributes
 * filter from which the output attributes are derived as
 * computeOutputAttributes(sources, requestedAttributes) derives them
 * (PotreeConverter.h:195-315: union of the sources' attributes, smallest
 * LAS scale, position first then the requested names that exist).  Each
 * chunk holds the first source's points, then the second's, ... (input
 * order inside a source).  Extra-bytes attributes that the filter or a mix
 * of point formats would move are refused: the reference writes them at
 * their offset in the source's layout (chunker_countsort_laszip.cpp:579-582),
 * outside their output attribute. */
int pcg_do_chunking_sources(const char* const* las_paths, int32_t num_sources, const char* outdir, const double* min,
                            const double* max, const char* const* attributes, int32_t num_attributes,
                            double* timings);

/* Both phases for several sources with the --attributes filter (main.cpp's
 * curateSources / computeOutputAttributes / computeStats, chunking and
 * indexing; the name in metadata.json is the first path's stem);
 * method, encoding, flags and timings as pcg_convert_encoded. */
int pcg_convert_sources(const char* const* las_paths, int32_t num_sources, const char* outdir, const char* method,
                        const char* encoding, const char* const* attributes, int32_t num_attributes, int32_t flags,
                        double* timings);

/* Phase 2 only (indexer::doIndexing, Converter/include/indexer.h:379) on an
 * existing outdir/chunks/ written by either chunker. */
int pcg_do_indexing(const char* outdir, const char* method, const char* name, int64_t points_total,
                    double* timings);

/* pcg_do_indexing with Options::encoding (as pcg_convert_encoded;
 * Converter/src/indexer.cpp:1453-1516). */
int pcg_do_indexing_encoded(const char* outdir, const char* method, const char* encoding, const char* name,
                            int64_t points_total, double* timings);

/* Per-cell counts of the chunker's counting pass
 * (countPointsInCells, Converter/src/chunker_countsort_laszip.cpp:131-361)
 * over n LAS records of bpp bytes at d_las.  d_grid receives grid_size^3
 * int32 counts (zeroed by the call).  min/max are the non-cubic bounds
 * handed to doChunking. */
int pcg_count_cells(const void* d_las, int64_t n, int32_t bpp, const double las_scale[3],
                    const double las_offset[3], const double pos_scale[3], const double pos_offset[3],
                    const double min[3], const double max[3], int32_t grid_size, int32_t* d_grid, void* stream);

/* pcg_count_cells without any allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, hosts that forbid hipMalloc).  The pass's
 * scratch (statistics, histogram, error flags) is a fixed 4 KiB device array
 * per GPU, held until the process ends, whatever n is; d_grid is cleared by a
 * kernel, not hipMemsetAsync.  Rejects bpp outside 12..128 and grid_size
 * outside 1..512.  Calls that overlap on one GPU share the array and race.
 * Same results and return codes as pcg_count_cells. */
int pcg_count_cells_noalloc(const void* d_las, int64_t n, int32_t bpp, const double las_scale[3],
                            const double las_offset[3], const double pos_scale[3], const double pos_offset[3],
                            const double min[3], const double max[3], int32_t grid_size, int32_t* d_grid,
                            void* stream);

/* Device-resident conversion for benchmarks and servers: the LAS records
 * are already in HBM (d_las, n records, header fields in `las_header`, a
 * 375-byte LAS 1.4 header image).  Produces the octree in HBM and, when
 * outdir is non-NULL, writes the three output files.  stats (nullable,
 * 24 doubles): [0] count s, [1] LUT+distribute s, [2] chunking s,
 * [3] indexing s, [4] octree build s, [5] sampling s, [6] octree assembly s,
 * [7] octree.bin bytes, [8] nodes, [9] sampled nodes, [10] candidates,
 * [11] FP64 re-checks, [12] FP32 pair tests, [13] max FP32 relative error
 * of re-checked pairs, [14] nodes sorted in libstdc++ order, [15] nodes
 * resolved sequentially, [16] chunks, [17] octree rounds, [18] depth,
 * [19] counting-pass kernel ms, [20] distribute kernels ms (HIP events),
 * [21] host tree set-up s, [22] host completion order s, [23] hierarchy.bin
 * and metadata.json s. */
int pcg_convert_device(const void* d_las, int64_t n, const uint8_t* las_header, const char* method,
                       const char* outdir, double* stats);

/* pcg_convert_device with Options::encoding (as pcg_convert_encoded).  With
 * "BROTLI" the octree image is assembled in host memory and stats[7] is its
 * compressed size. */
int pcg_convert_device_encoded(const void* d_las, int64_t n, const uint8_t* las_header, const char* method,
                               const char* encoding, const char* outdir, double* stats);

/* Phase 2 on chunk records already in HBM: indexer::doIndexing
 * (Converter/include/indexer.h:379) over the chunks that getChunks
 * (Converter/src/indexer.cpp:53-179) would read from a chunks/ directory.
 * d_records holds the records of num_chunks chunks back to back, in
 * ascending chunk name ("r" + octal digits, the names of the <name>.bin
 * files), chunk_counts[i] records each; chunks_metadata is the text of
 * chunks/metadata.json (cube min/max, attributes, scale, offset).
 * d_records is only read (the library indexes a copy).  method, encoding and
 * name as pcg_do_indexing_encoded; outdir (nullable) receives octree.bin,
 * hierarchy.bin and metadata.json; stats (nullable) as pcg_convert_device,
 * with the chunking entries 0. */
int pcg_do_indexing_device(const void* d_records, int32_t num_chunks, const char* const* chunk_names,
                           const int64_t* chunk_counts, const char* chunks_metadata, const char* method,
                           const char* encoding, const char* name, const char* outdir, double* stats);

/* One Sampler::sample call (Converter/include/structures.h:152-164) of
 * SamplerPoisson (sampler_poisson.h:14-304), SamplerPoissonAverage
 * (sampler_poisson_average.h:14-492) or SamplerRandom (sampler_random.h),
 * by method "poisson" / "poisson_average" / "random", on the GPU: the
 * plugin a Sampler subclass forwards to (its glue is in INTEGRATION.md).
 * The caller flattens the subtree below the node handed to sample() into
 * num_nodes nodes, node 0 being that node: children (8 ids per node, -1 =
 * none, every other node below node 0 exactly once), bounds (6 doubles per
 * node: min x, y, z, max x, y, z), levels (name.size() - 1; a child one
 * below its parent), sampled (nullable; 1 for a node sampled before, which
 * the traversal does not enter: it only lends its records), counts and
 * records: the records (bpp bytes each, positions as int32 X, Y, Z at
 * byte 0) of the nodes without children, back to back in node order; nodes
 * with children hold none.  rgb_offset locates the uint16 rgb triple
 * (Poisson-average); scale / offset are Attributes::posScale / posOffset;
 * base_spacing is the sampler's baseSpacing.  Random leaf shuffles take
 * the "random_seed" option.
 * Results: out_counts[i] and, back to back in node order in out_records
 * (capacity out_capacity records), every node's records after the call;
 * out_children (8 per node) with the children the sampler unlinked set to
 * -1; events (2 int32 per callback: node, 0 = onNodeCompleted,
 * 1 = onNodeDiscarded; at most 2 * num_nodes callbacks) in the order the
 * reference calls them, and *num_events.  The records of a node unlinked
 * without a callback (Poisson-average) are unspecified.  Poisson and
 * random move records only (out_capacity >= sum(counts) suffices);
 * Poisson-average adds each sampled node's accepted copies
 * (sum(counts) * (levels spanned + 1) suffices); a smaller capacity fails
 * with the needed size in the error text. */
int pcg_sample_subtree(const char* method, int32_t num_nodes, const int32_t* children, const double* bounds,
                       const int32_t* levels, const uint8_t* sampled, const int64_t* counts, const void* records,
                       int32_t bpp, int32_t rgb_offset, const double scale[3], const double offset[3],
                       double base_spacing, int64_t* out_counts, void* out_records, int64_t out_capacity,
                       int32_t* out_children, int32_t* events, int32_t* num_events);

/* Device memory helpers for callers without a HIP runtime binding.
 * kind: 0 host->device, 1 device->host, 2 device->device. */
int pcg_device_alloc(void** ptr, uint64_t bytes);
int pcg_device_free(void* ptr);
int pcg_copy(void* dst, const void* src, uint64_t bytes, int32_t kind);
int pcg_synchronize(void);
/* Free and total device memory (hipMemGetInfo) and the bytes the library
 * holds in its reusable buffers (none of it is freed between jobs unless
 * pcg_release is called or a job needs the room). */
int pcg_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);
uint64_t pcg_workspace_bytes(void);

/* Synthetic LAS generator (not part of the reference): preset 0 = 500 m
 * tile, 1 = 5 km tile in UTM-like coordinates, 2 = precision scene (scale
 * 0.0001, offset 1e6 m, near-duplicate clusters, a lattice whose pitch is a
 * node spacing).  `params` points at a
 * caller-owned buffer of at least 256 bytes. */
int pcg_lasgen_preset(int32_t preset, int64_t n, uint64_t seed, void* params);
/* Record format of the generated points (0-7, presets use 2) and whether 7
 * extra bytes (float32, uint16, int8) follow each record. */
int pcg_lasgen_configure(void* params, int32_t point_format, int32_t extra_bytes);
int32_t pcg_lasgen_record_length(const void* params);
void pcg_lasgen_fill(const void* params, int64_t first, int64_t count, uint8_t* out);
int pcg_lasgen_write(const void* params, const char* path);
/* Generates records [first, first+count) into d_out on the GPU and folds the
 * integer X/Y/Z extents into d_qminmax[6] (min xyz, max xyz). */
int pcg_lasgen_device(const void* params, int64_t first, int64_t count, uint8_t* d_out, int32_t* d_qminmax);
/* 375-byte LAS 1.4 header image for generated points with the given
 * integer extents. */
int pcg_lasgen_header(const void* params, const int32_t qminmax[6], uint8_t* header375);

/* Synthetic pre-chunked input of phase 2 (not part of the reference): the
 * cube [0, 2048 m)^3 at scale 0.001, num_chunks (a power of two up to 4096)
 * chunks "rABCD" of octree level 4 spread over its 16^3 cells, each a local
 * terrain patch with box buildings, records in the output layout of LAS
 * point format 2 plus three extra-bytes attributes (float32 N(0,1) "gauss",
 * uint16 "uniform16", int8 "uniform8"; 34 bytes).  Chunk point counts are
 * log-normal: round(median * exp(sigma * g)), g ~ N(0,1), clipped to
 * [min_count, max_count].  `params` points at a caller-owned buffer of at
 * least 128 bytes. */
int pcg_chunkgen_params(int32_t num_chunks, double median, double sigma, int64_t min_count, int64_t max_count,
                        uint64_t seed, void* params);
/* counts (nullable, num_chunks values) and names (nullable, 6 bytes each:
 * "rABCD" and a terminator) of the chunks in ascending name order. */
int pcg_chunkgen_counts(const void* params, int64_t* counts, char* names);
/* records [first, first+count) of chunk `chunk` into out (host) */
void pcg_chunkgen_fill(const void* params, int32_t chunk, int64_t first, int64_t count, uint8_t* out);
/* Writes outdir/chunks/<name>.bin and outdir/chunks/metadata.json as the
 * reference chunker would have left them (host generation). */
int pcg_chunkgen_write(const void* params, const char* outdir);
/* Generates every chunk's records into d_records on the GPU (34 bytes per
 * point, ascending name order) and writes the text of the matching
 * chunks/metadata.json (attribute minima, maxima and the classification
 * histogram of the generated records) into metadata. */
int pcg_chunkgen_device(const void* params, void* d_records, char* metadata, int64_t metadata_capacity);

#ifdef __cplusplus
}
#endif

#endif /* POTREE_GPU_H */
