// Launchers for the gfx950 chunker kernels (chunker.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lasgen.h"
#include "pcg_common.h"

namespace pcg {

int chunker_set_params(const ChunkerParams& p, hipStream_t stream);
// d_cells (n entries, may be null) receives the grid cell of every point
// (0xFFFFFFFF outside the box)
// Counting pass without any allocation: k_count's statistics, histogram and
// error flags live in a fixed device array (4 KiB per GPU, whatever n is),
// cleared by fill kernels before the pass; *h_err receives the error flags.
// Calls that overlap on one GPU share that array and race.
int chunker_count_noalloc(const uint8_t* d_las, int64_t n, int32_t* d_grid, int64_t cells, int num_cus,
                          hipStream_t stream, int32_t* h_err);
// Device address of that array on the current device (looked up once per device).
int noalloc_scratch(uint8_t** out);

int chunker_count(const uint8_t* d_las, int64_t n, int32_t* d_grid, uint32_t* d_cells, uint64_t* d_stat_min,
                  uint64_t* d_stat_max, unsigned long long* d_class_hist, int32_t* d_err, int num_cus,
                  hipStream_t stream);
// d_levels[l] holds (2^l)^3 int32 values; d_levels[level_max] is the count grid.
int chunker_pyramid(int32_t* const* d_levels, int32_t level_max, int64_t cap, hipStream_t stream);
int chunker_chunk_cells(int32_t* const* d_levels, int32_t level_max, int64_t* d_out, uint32_t* d_counter,
                        uint32_t max_out, hipStream_t stream);
int chunker_lut_fill(const int64_t* d_chunks, int64_t num_chunks, int32_t level_max, int32_t* d_lut,
                     hipStream_t stream);
// d_ids holds the cells written by chunker_count and is overwritten with
// the chunk index of every point
int chunker_distribute(const uint8_t* d_las, uint32_t* d_ids, int64_t n, const int32_t* d_lut, int32_t num_chunks,
                       int32_t tile_size, uint32_t* d_M, void* d_scan_ws, uint8_t* d_out, int32_t* d_err,
                       hipStream_t stream);
// Distribute for the record layouts of point formats 0-3 and 6-7 (records
// up to 64 bytes, fewer than 65535 chunks): the tile histogram writes every
// record's chunk as 16 bits (d_ids, n entries) and the tile matrix, the scan
// gives every (chunk, tile) its first record, and the tile transcode writes
// the records with aligned 16-byte stores.  Tiles of 4096 records; d_M and
// d_scan_ws as for chunker_distribute.
bool distribute_tail_fits(int32_t num_chunks);
int chunker_distribute_tail(const uint8_t* d_las, const uint32_t* d_cells, uint16_t* d_ids, int64_t n,
                            const int32_t* d_lut, int32_t num_chunks, uint32_t* d_M, void* d_scan_ws,
                            uint8_t* d_out, int32_t* d_err, hipStream_t stream);
// The two halves of a distribute pass, per source: the tile histogram of
// one source's points (its tiles are columns tile_base .. of the chunk x
// tile matrix d_M, whose rows hold `stride` tiles; with d_ids16 the chunk
// ids go there as 16 bits, else over the cells) and, after the scan of the
// whole matrix, the placement of that source's records (the tile transcode
// with d_ids16, the byte-wise scatter without).  Tiles of 4096 points.
int chunker_tile_hist(uint32_t* d_cells, uint16_t* d_ids16, int64_t n, const int32_t* d_lut, int32_t num_chunks,
                      uint32_t* d_M, int64_t tile_base, int64_t stride, int32_t* d_err, hipStream_t stream);
int chunker_place(const uint8_t* d_las, const uint32_t* d_ids32, const uint16_t* d_ids16, int64_t n,
                  int32_t num_chunks, const uint32_t* d_M, int64_t tile_base, int64_t stride, uint8_t* d_out,
                  hipStream_t stream);
int lasgen_device(const LasGenParams& p, int64_t first, int64_t count, uint8_t* d_out, int32_t* d_qminmax,
                  hipStream_t stream);

}  // namespace pcg
