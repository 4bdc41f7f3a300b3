// ofp_internal.h -- shared definitions between the HIP kernels and the host
// runtime of libofp.so (not part of the public ABI; see include/ofp.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ofp {

constexpr int kMaxW = 28;          // W = n+1 <= 28
constexpr int kDpRows = kMaxW + 2; // positions 0..kMaxW+1
constexpr int kDpCols = 32;        // heights 0..31 (h+2 <= W+2 <= 30)
constexpr int kTile = 64;          // AoSoA tile: 64 states x K moduli
constexpr int kMaxKC = 16;         // moduli handled by one launch

// Motzkin-height table: D[w][h] = number of words over w positions entered at
// height h that never go negative and end at height 0.  Same recurrence as the
// reference's H_DP (src/cuda/b300/oneesan_cuda_gridfp_b300_hbm32_fullmate_dropN_batch.cu:62).
void build_height_table(uint64_t (*d)[kDpCols]);

// Element (state i, modulus j) of an interleaved state array with K moduli:
// tiles of 64 consecutive states; inside a tile, modulus-major.  A wave of 64
// lanes reading one modulus of 64 consecutive states reads 256 contiguous
// bytes; all K residues of a state stay inside one 256*K-byte tile.
__host__ __device__ inline uint64_t aosoa(uint64_t i, int K, int j) {
    return (i >> 6) * (uint64_t)(kTile * K) + (uint64_t)j * kTile + (i & (kTile - 1));
}
inline uint64_t padded_states(uint64_t n) { return (n + kTile - 1) / kTile * kTile; }

struct CellLaunch {
    uint32_t* main;       // interleaved main residues
    uint32_t* blk;        // interleaved blocked residues (read and written in place)
    const void* dp;       // device height table (uint32 or uint64 entries, kDpRows x kDpCols)
    uint64_t nblk;        // blocked words processed: [first, first + nblk)
    uint64_t first;       // first blocked word index (0 for a whole sweep)
    int W, p;             // frontier width, cell
    int read_blk;         // blocked input of this cell is valid
    int K, off, kc;       // interleave width, first modulus, moduli in this launch
    uint32_t mod[kMaxKC];
};

// Launch one cell step (all blocked words, moduli off..off+kc-1).
hipError_t launch_cell(const CellLaunch& a, bool wide_index, int threads, int blocks,
                       hipStream_t s);
// Set main[aosoa(index,K,j)] = 1 mod moduli[j] for j < K.
hipError_t launch_set_one(uint32_t* main, uint64_t index, int K, const uint32_t* mods_dev,
                          hipStream_t s);

}  // namespace ofp
