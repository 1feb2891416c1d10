// Launch geometry chosen by cytom_amd/autotune.py per (scale, edge factor)
// of the RMAT update stream.  When vertices_per_tile is 0, cytom_create takes
// the rows of the edge factor nearest the configuration's (edge_capacity / V,
// or 16 without a capacity) and among them the row of the largest scale not
// above log2(V), or the smallest scale when every row is above it.  Generated
// from cytom_amd/tuning_table.json by `python -m cytom_amd.autotune --install`.
#pragma once

#include <cstdint>

namespace cytom {

struct TuneRow {
  uint32_t log2v, edge_factor, vpt, chunk_slots, source_bands;
};

constexpr TuneRow TUNE_TABLE[] = {
    {16, 16, 4096, 256, 4},  // 13.3 GTEPS
    {18, 16, 4096, 256, 1},  // 41.3 GTEPS
    {20, 16, 4096, 256, 1},  // 129.2 GTEPS
    {22, 16, 16384, 256, 1},  // 208.9 GTEPS
    {24, 16, 16384, 256, 4},  // 194.6 GTEPS
};

}  // namespace cytom
