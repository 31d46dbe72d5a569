// cell_sort.h -- counting sort of the particles by cell key over the dense
// cell table; replaces thrust::sort_by_key + compute_cell_start_end
// (particle_simulation.cu:2340-2343, 2362-2368, 155-193) when the key space
// is a small multiple of the particle count.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

// Largest key space (cells per particle) for which the counting sort is
// used: its two reads of the count array cost 8 bytes per cell against the
// radix sort's ~48 bytes per particle.
constexpr uint32_t kCellSortRatio = 16;

struct CellSort {
    uint32_t cells = 0;       // key space (global cell count)
    uint32_t blocks = 0;      // cell blocks of the scan
    uint32_t* cnt = nullptr;  // [cells] particles per cell; all zero between sorts
    uint32_t* rank = nullptr; // [capacity] arrival rank of each particle in its cell
    uint32_t* bsum = nullptr; // [blocks] particles per 4096-cell block
    uint32_t* used = nullptr; // [blocks] nonzero where a key of this sort falls; all zero between sorts
    size_t bytes = 0;

    hipError_t init(uint32_t cells, uint32_t capacity);
    void release();
    // keys: cell key of each storage slot (n of them).  Writes skey/perm (the
    // slots ordered by key; inside a cell in arrival order, which the
    // reorder pass replaces by its own tie order) and range[k] = [start,
    // end) for every occupied cell k.  range entries of other cells are not
    // touched, except that clear_prev = true first invalidates the cells
    // of the previous sort of this key array (skey as it holds on entry,
    // n keys); otherwise the caller does that.
    void sort(hipStream_t st, const uint32_t* keys, uint32_t n, uint32_t* skey, uint32_t* perm, int2* range,
              bool clear_prev);
    // The same with the count read on the device (n_dev <= capacity), so
    // that the call can be captured in a graph; grids sized for capacity.
    void sort(hipStream_t st, const uint32_t* keys, const uint32_t* n_dev, uint32_t capacity, uint32_t* skey,
              uint32_t* perm, int2* range);

  private:
    void launch(hipStream_t st, const uint32_t* keys, uint32_t n, const uint32_t* n_dev, uint32_t grid_n,
                uint32_t* skey, uint32_t* perm, int2* range, const uint32_t* prev_skey, uint32_t prev_n);
};

}  // namespace sph
