// radix_sort.h -- LDS-staged stable LSD radix sort of (uint32 key, uint32
// value) pairs, 8-bit digits.  Replaces thrust::sort_by_key
// (particle_simulation.cu:2340-2343, 2350-2353).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

constexpr int kSortThreads = 256;   // 4 wavefronts of 64
constexpr int kSortItems = 8;       // keys per thread per tile
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kSortBins = 256;

struct RadixSort {
    uint32_t capacity = 0;
    uint32_t max_tiles = 0;
    uint32_t* key_alt = nullptr;
    uint32_t* val_alt = nullptr;
    uint32_t* hist = nullptr;     // [256][max_tiles] per-tile counts -> row prefix
    uint32_t* totals = nullptr;   // [256] digit totals

    hipError_t init(uint32_t cap);
    void release();
    size_t bytes() const;
    // Sorts n = *n_dev pairs by the low `bits` bits of the key (stable).
    // vals_in == nullptr means values 0..n-1.  Outputs must not alias the
    // inputs.  The element count is read on the device so the call can be
    // captured in a graph; grids are sized for `capacity`.
    void sort(hipStream_t st, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
              uint32_t* vals_out, const uint32_t* n_dev, int bits);
};

}  // namespace sph
