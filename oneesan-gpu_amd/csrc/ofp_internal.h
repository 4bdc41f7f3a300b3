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
constexpr int kMaxCap = 15;        // height caps 1..kMaxCap have packed tables (max height at W=28 is 15)

// Packed height table (cell::Quad entries) for device upload; returns bytes.
size_t pack_height_table(bool wide, void* out, size_t cap);

// Motzkin-height table: D[w][h] = number of words over w positions entered at
// height h that never go negative and end at height 0.  Same recurrence as the
// reference's H_DP (src/cuda/b300/oneesan_cuda_gridfp_b300_hbm32_fullmate_dropN_batch.cu:62).
void build_height_table(uint64_t (*d)[kDpCols]);
// Height table of words whose every height (entry height included) is <= cap:
// d[w][h] = number of words of width w entered at height h <= cap that end at 0
// without leaving [0, cap].
void build_capped_table(int cap, uint64_t (*d)[kDpCols]);
// Occupancy group `mask` of a row of width W split at `low` (cell::group_occupancy):
// a LOW group (high = false; fixed positions 0..low-1, cells W-1..low+1) or a
// HIGH group (fixed positions low+1..W-1, cells low..1).  Its packed 32-bit
// table (kDpRows x kDpCols cell::Quad<uint32_t>; returns the bytes written, 0 if
// cap is too small) and its main / blocked word counts.
size_t pack_group_table(bool high, int W, int low, uint32_t mask, void* out, size_t cap);
void group_counts(bool high, int W, int low, uint32_t mask, uint64_t* main_words, uint64_t* blocked_words);
// Packed capped tables for caps 1..kMaxCap (cell::Pair<uint32 or uint64>,
// kMaxCap x kDpRows x kDpCols) into out; returns the bytes written, 0 if cap is too small.
size_t pack_capped_tables(bool wide, void* out, size_t cap);

// Element (state i, modulus j) of an interleaved state array with K moduli:
// tiles of 64 consecutive states; inside a tile, modulus-major.  A wave of 64
// lanes reading one modulus of 64 consecutive states reads 256 contiguous
// bytes; all K residues of a state stay inside one 256*K-byte tile.
__host__ __device__ inline uint64_t aosoa(uint64_t i, int K, int j) {
    return (i >> 6) * (uint64_t)(kTile * K) + (uint64_t)j * kTile + (i & (kTile - 1));
}
inline uint64_t padded_states(uint64_t n) { return (n + kTile - 1) / kTile * kTile; }
// Element (i, j) under either state layout (0 = 64-state tiles, 1 = state-major).
__host__ __device__ inline uint64_t state_elem(int layout, uint64_t i, int K, int j) {
    return layout == 0 ? aosoa(i, K, j) : i * (uint64_t)K + (uint64_t)j;
}
// Interleave stride for k moduli: state-major layout pads to an even stride so
// that two residues move per 8-byte access.
inline int interleave_stride(int layout, int k) { return (layout == 1 && k > 1) ? (k + 1) / 2 * 2 : k; }

struct CellLaunch {
    uint32_t* main;       // interleaved main residues
    uint32_t* blk;        // interleaved blocked residues (read and written in place)
    const void* dp;       // device packed height table (cell::Quad<uint32 or uint64>, kDpRows x kDpCols)
    uint64_t nblk;        // blocked words processed: [first, first + nblk)
    uint64_t first;       // first blocked word index (0 for a whole sweep)
    int W, p;             // frontier width, cell
    int read_blk;         // blocked input of this cell is valid
    int K, off, kc;       // interleave stride, first modulus, moduli in this launch
    int layout;           // 0: 64-state tiles (AoSoA), 1: state-major (AoS)
    int variant;          // state-major only: cell_kernel V parameter (0..3)
    const void* dpc;      // capped launches only: packed capped table of this launch's cap
    int xcd_shift;        // chunks are dealt to the 8 XCDs in pieces of 2^xcd_shift (< 0: eighths)
    uint32_t mod[kMaxKC];
    // shard launches only (launch_cell_shard): resident main window and remote seed
    uint64_t m0, M;
    uint64_t seed;
    unsigned long long* stray;
};

// Launch one cell step (all blocked words, moduli off..off+kc-1).
hipError_t launch_cell(const CellLaunch& a, bool wide_index, int threads, int blocks,
                       hipStream_t s);
// Launch one cell over only the blocked words whose heights stay <= the cap of
// a.dpc (nblk = their number); the other blocked words are skipped.
hipError_t launch_cell_capped(const CellLaunch& a, bool wide_index, int threads, int blocks,
                              hipStream_t s);
// Launch one cell over the blocked words [first, first+nblk) of a shard whose
// main window is [m0, m0+M) (state-major layout; main/blk point at the window
// arrays).  Main words outside the window read as remote_residue(seed, ...).
hipError_t launch_cell_shard(const CellLaunch& a, bool wide_index, int threads, int blocks,
                             hipStream_t s);
// Fill a shard window with remote_residue values: main words [m0, m0+M) and
// blocked words [b0, b0+B) (blocked indices in their own hash domain).
hipError_t launch_fill_remote(uint32_t* main, uint64_t m0, uint64_t M, uint32_t* blk, uint64_t b0,
                              uint64_t B, int K, int k, const uint32_t* mods_dev, uint64_t seed,
                              hipStream_t s);
// Set main[(index, j)] = 1 mod moduli[j] for j < k (interleave stride K).
hipError_t launch_set_one(uint32_t* main, uint64_t index, int K, int k, int layout,
                          const uint32_t* mods_dev, hipStream_t s);
// The same with the moduli passed by value from host memory (k <= kMaxKC).
hipError_t launch_set_one_values(uint32_t* main, uint64_t index, int K, int k, int layout,
                                 const uint32_t* mods_host, hipStream_t s);

// ---- pair pass (ofp_pair.hip; launch struct pair::PairLaunch in ofp_pair_impl.h) ----
namespace pair { struct PairLaunch; }
// Packed class tables for every (p, d), p in 0..kMaxW-1, d in -3..3, each
// class_table_bytes(wide) long at offset (p*7 + d+3) * class_table_bytes(wide).
size_t pack_class_tables(bool wide, void* out, size_t cap);
size_t class_table_bytes(bool wide);
// Classes of (p, d) in a row of width W.
uint64_t pair_class_count(int W, int p, int d);
// One pair-pass launch (a.kc in {1, 2, 4, ..., 16} moduli, 32-bit ranks).
hipError_t launch_pair(const pair::PairLaunch& a, bool wide_index, int threads, int blocks, hipStream_t s);

// ---- exchange between the two placements of a row (ofp_exchange.hip) ----
// A tile: `len` consecutive words of one group of a LOW window (local ranks
// rank0.., window index word0..).
struct XTile {
    uint64_t word0;
    uint32_t group, rank0, len, pad;
};
constexpr int kMaxShares = 16;
constexpr uint32_t kXTileWords = 32768;  // words per tile (at most)
struct XchgLaunch {
    const uint32_t* res;     // the window's residues of this kind (main or blocked), stride K
    const char* gtab;        // the window's packed 32-bit group tables (kDpRows x kDpCols quads each)
    const XTile* tiles;
    uint32_t ntiles;
    int K;                   // interleave stride: a record carries K residues
    int top, base, nf;       // top position of the kind, first fixed HIGH position, fixed positions
    int main;                // 1: main words (width W), 0: blocked words (width W-1)
    const uint32_t* ht;      // group-row table of this kind (ofp_exchange_impl.h)
    const void* dplain;      // plain packed 32-bit height table
    const uint64_t* base_of; // per HIGH mask: offset of its group in its share's placement
    const uint32_t* lut;     // bottoms' plain-numbering rank contributions (xchg::build_bottom_lut), all groups
    const uint32_t* lut_off; // per group of the window: kDpCols + 1 offsets into lut, by entry height
    uint32_t first[kMaxShares + 1];  // share s owns HIGH masks first[s]..first[s+1]-1
    int shares;
    uint32_t* counts;        // [shares][ntiles] words per destination share and tile
    uint64_t* offs;          // [shares][ntiles] exclusive prefix within the share
    uint64_t* totals;        // [shares] records per share
    uint64_t* out_idx;       // records: destination index in the share's HIGH placement
    uint32_t* out_res;       //          and the K residues; bucket s after buckets 0..s-1
};
// Count, scan and write launches of one kind (in that order on one stream).
hipError_t launch_exchange_count(const XchgLaunch& a, hipStream_t s);
hipError_t launch_exchange_scan(const XchgLaunch& a, hipStream_t s);
hipError_t launch_exchange_write(const XchgLaunch& a, hipStream_t s);

}  // namespace ofp
