// ofp_exchange.hip -- the exchange between the two placements of a row, for
// gfx950 (index math in ofp_exchange_impl.h).
//
// After the high half of a row has run over a LOW window (ofp_group_sweep),
// every word of the window moves to its HIGH group, owned by one of `shares`
// GPUs.  This file writes the window's words into staging buckets, one per
// destination share, each word as a record (its index in the destination
// share's HIGH placement, its K residues): the send side of the per-row
// all-to-all (reference: the 8x8 bucket transpose between the H-major and
// L-major placements, docs/b300-bucket-snake.md:7-20).
//
// Three launches per kind of word (main, blocked), one workgroup per tile of
// at most kXTileWords words of one group.  Words with one top part (the fixed
// HIGH positions) are a run of consecutive ranks -- a block, about a hundred
// words at W = 28 -- with one destination share and one top-part HIGH-rank
// contribution, so each thread walks a run of consecutive ranks block by
// block, decoding one top part per block it meets (xchg::decode_top), and takes
// each word's bottom contribution from the group's bottom table
// (xchg::build_bottom_lut):
//   count  -- words per destination share and tile (each thread walks a run
//             of the tile);
//   scan   -- one workgroup per share: exclusive prefix of the tile counts;
//   write  -- write each word's record.  Records keep the window order inside a
//             bucket: each round of 2048 words is compacted with wave64 ballots
//             (a lane's slot = the tile's cursor + the set lanes below it + the
//             counts of the lower waves and slices), so the buckets are
//             deterministic.
// The residues are read once in window order (coalesced) and written once as
// records; records of one share in one round are contiguous.
#include "ofp_exchange_impl.h"

namespace ofp {

namespace {

using cell::Quad;
constexpr int kXThreads = 256;
constexpr int kXWaves = kXThreads / 64;
constexpr int kXSlices = 8;  // slices of 256 words per round of the write pass

__device__ void load_quads(Quad<uint32_t> (*dst)[kDpCols], const Quad<uint32_t>* src) {
    Quad<uint32_t>* d = &dst[0][0];
    for (int t = threadIdx.x; t < kDpRows * kDpCols; t += blockDim.x) d[t] = src[t];
}

__device__ const Quad<uint32_t>* group_table(const XchgLaunch& a, uint32_t g) {
    return reinterpret_cast<const Quad<uint32_t>*>(a.gtab + (size_t)g * sizeof(Quad<uint32_t>) * kDpRows * kDpCols);
}

// The group's bottom offsets (kDpCols + 1 entries) into LDS.
__device__ void load_off(uint32_t* dst, const XchgLaunch& a, uint32_t g) {
    const uint32_t* src = a.lut_off + (size_t)g * (kDpCols + 1);
    for (int t = threadIdx.x; t < kDpCols + 1; t += blockDim.x) dst[t] = src[t];
}

// Count pass: thread t takes a contiguous run of the tile's ranks and walks it
// block by block -- one top-part decode per block it meets (about one per
// thread: blocks are longer than the runs), adding the block's words in the
// run to the block's destination share.
template <bool Main>
__device__ void count_run(const XchgLaunch& a, const XTile& t, const Quad<uint32_t> (*tq)[kDpCols],
                          const uint32_t* off, uint32_t* cnt) {
    const uint32_t per = (t.len + kXThreads - 1) / kXThreads;
    uint32_t i = min(threadIdx.x * per, t.len);
    const uint32_t e = min(i + per, t.len);
    while (i < e) {
        const xchg::Top tp = xchg::decode_top<Main>(tq, a.top, a.base, t.rank0 + i);
        // the block holds ranks r - r_rest .. r - r_rest + len - 1, r_rest < len
        const uint32_t end = i + (off[tp.h + 1] - off[tp.h] - tp.r_rest);
        const uint32_t stop = min(end, e);
        atomicAdd(&cnt[xchg::share_of(tp.mask, a.first, a.shares)], stop - i);
        i = stop;
    }
}

__global__ __launch_bounds__(kXThreads) void xchg_count_kernel(XchgLaunch a) {
    __shared__ Quad<uint32_t> tq[kDpRows][kDpCols];
    __shared__ uint32_t cnt[kMaxShares];
    __shared__ uint32_t off[kDpCols + 1];
    const XTile t = a.tiles[blockIdx.x];
    load_quads(tq, group_table(a, t.group));
    load_off(off, a, t.group);
    if (threadIdx.x < kMaxShares) cnt[threadIdx.x] = 0;
    __syncthreads();
    if (a.main) count_run<true>(a, t, tq, off, cnt);
    else count_run<false>(a, t, tq, off, cnt);
    __syncthreads();
    if ((int)threadIdx.x < a.shares) a.counts[(size_t)threadIdx.x * a.ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// One workgroup per share: offs = exclusive prefix of that share's tile counts.
__global__ __launch_bounds__(1024) void xchg_scan_kernel(XchgLaunch a) {
    __shared__ uint64_t part[1024];
    const int s = blockIdx.x, tid = threadIdx.x;
    const uint32_t* c = a.counts + (size_t)s * a.ntiles;
    uint64_t* o = a.offs + (size_t)s * a.ntiles;
    const uint32_t per = (a.ntiles + 1023u) / 1024u;
    const uint32_t b = (uint32_t)tid * per, e = b + per < a.ntiles ? b + per : a.ntiles;
    uint64_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += c[i];
    part[tid] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const uint64_t v = tid >= d ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint64_t run = part[tid] - sum;
    for (uint32_t i = b; i < e; ++i) {
        o[i] = run;
        run += c[i];
    }
    if (tid == 1023) a.totals[s] = part[1023];
}

// The records of a run of kXSlices consecutive words of a round, by a block
// walk (one top-part decode per block met, about one per run): destination
// share (-1 past the tile) and index, into the round's LDS arrays.
template <bool Main>
__device__ void walk_run(const XchgLaunch& a, const XTile& t, const Quad<uint32_t> (*tq)[kDpCols],
                         const uint32_t* off, uint32_t r0, uint64_t* wdst, int8_t* wsh) {
    const uint32_t first = r0 + (uint32_t)(threadIdx.x * kXSlices);
    const uint32_t e = min(first + (uint32_t)kXSlices, t.len);
    for (uint32_t x = e > first ? e : first; x < first + (uint32_t)kXSlices; ++x) wsh[x - r0] = -1;
    uint32_t i = first;
    while (i < e) {
        const xchg::Top tp = xchg::decode_top<Main>(tq, a.top, a.base, t.rank0 + i);
        const uint32_t lut0 = off[tp.h] + tp.r_rest;  // bottom table entry of rank i
        const uint32_t stop = min(i + (off[tp.h + 1] - off[tp.h] - tp.r_rest), e);
        const uint64_t base = a.base_of[tp.mask];
        const uint32_t top = xchg::top_high(tp, a.ht);
        const int8_t sh = (int8_t)xchg::share_of(tp.mask, a.first, a.shares);
        for (uint32_t x = i; x < stop; ++x) {
            wdst[x - r0] = base + (uint32_t)(top + a.lut[lut0 + (x - i)]);
            wsh[x - r0] = sh;
        }
        i = stop;
    }
}

// Rounds of R = kXSlices x 256 words.  Each thread loads the residues of its
// words of the round's slices (slice j = the 256 consecutive ranks r0 + 256 j
// ..; kXSlices loads per thread in flight), walks the blocks of its run of
// kXSlices consecutive words (walk_run) into LDS, then, slice by slice, counts
// the records per destination share and wave; one thread per share turns the
// counts of the round's (slice, wave) pairs, in window order, into absolute
// bucket slots, and the words are stored.
__global__ __launch_bounds__(kXThreads) void xchg_write_kernel(XchgLaunch a) {
    constexpr int E = kXSlices * kXWaves;  // (slice, wave) pairs of a round, in window order
    constexpr int R = kXSlices * kXThreads;
    __shared__ Quad<uint32_t> tq[kDpRows][kDpCols];
    __shared__ uint64_t cursor[kMaxShares];
    __shared__ uint32_t wcnt[E][kMaxShares];
    __shared__ uint64_t sbase[E][kMaxShares];
    __shared__ uint32_t off[kDpCols + 1];
    __shared__ uint64_t wdst[R];
    __shared__ int8_t wsh[R];
    const XTile t = a.tiles[blockIdx.x];
    load_quads(tq, group_table(a, t.group));
    load_off(off, a, t.group);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < a.shares) {
        uint64_t start = 0;
        for (int q = 0; q < tid; ++q) start += a.totals[q];
        cursor[tid] = start + a.offs[(size_t)tid * a.ntiles + blockIdx.x];
    }
    __syncthreads();  // the tables and cursors are in LDS before any thread decodes
    const uint64_t below = (1ull << lane) - 1ull;
    const bool even = (a.K & 1) == 0;
    for (uint32_t r0 = 0; r0 < t.len; r0 += R) {
        uint2 v01[kXSlices], v23[kXSlices];
#pragma unroll
        for (int j = 0; j < kXSlices; ++j) {
            const uint32_t i = r0 + (uint32_t)(j * kXThreads + tid);
            v01[j] = make_uint2(0u, 0u);
            v23[j] = make_uint2(0u, 0u);
            if (i < t.len && even) {
                const uint32_t* src = a.res + (t.word0 + i) * (uint64_t)a.K;
                v01[j] = *reinterpret_cast<const uint2*>(src);
                if (a.K >= 4) v23[j] = *reinterpret_cast<const uint2*>(src + 2);
            }
        }
        if (a.main) walk_run<true>(a, t, tq, off, r0, wdst, wsh);
        else walk_run<false>(a, t, tq, off, r0, wdst, wsh);
        __syncthreads();
        int sh[kXSlices];
        uint32_t pre[kXSlices];
#pragma unroll
        for (int j = 0; j < kXSlices; ++j) {
            sh[j] = wsh[j * kXThreads + tid];
            const bool valid = sh[j] >= 0;
            const int e = j * kXWaves + wave;
            const uint64_t vm = __ballot(valid);
            const int s0 = vm ? __shfl(sh[j], __ffsll((long long)vm) - 1) : 0;
            if (__ballot(valid && sh[j] != s0) == 0) {
                // one destination share for the wave's words of this slice
                pre[j] = (uint32_t)__popcll(vm & below);
                if (lane < a.shares) wcnt[e][lane] = lane == s0 ? (uint32_t)__popcll(vm) : 0u;
            } else {
                pre[j] = 0;
                for (int q = 0; q < a.shares; ++q) {
                    const uint64_t m = __ballot(sh[j] == q);
                    if (sh[j] == q) pre[j] = (uint32_t)__popcll(m & below);
                    if (lane == 0) wcnt[e][q] = (uint32_t)__popcll(m);
                }
            }
        }
        __syncthreads();
        if (tid < a.shares) {
            uint64_t run = cursor[tid];
            for (int e = 0; e < E; ++e) {
                sbase[e][tid] = run;
                run += wcnt[e][tid];
            }
            cursor[tid] = run;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kXSlices; ++j) {
            if (sh[j] < 0) continue;
            const uint32_t i = r0 + (uint32_t)(j * kXThreads + tid);
            const uint64_t slot = sbase[j * kXWaves + wave][sh[j]] + pre[j];
            // records are written once and not read here again: non-temporal stores
            __builtin_nontemporal_store(wdst[j * kXThreads + tid], &a.out_idx[slot]);
            const uint32_t* src = a.res + (t.word0 + i) * (uint64_t)a.K;
            uint32_t* out = a.out_res + slot * (uint64_t)a.K;
            if (even) {
                __builtin_nontemporal_store(v01[j].x, out);
                __builtin_nontemporal_store(v01[j].y, out + 1);
                if (a.K >= 4) {
                    __builtin_nontemporal_store(v23[j].x, out + 2);
                    __builtin_nontemporal_store(v23[j].y, out + 3);
                }
                for (int k = 4; k < a.K; k += 2)
                    *reinterpret_cast<uint2*>(out + k) = *reinterpret_cast<const uint2*>(src + k);
            } else {
                for (int k = 0; k < a.K; ++k) out[k] = src[k];
            }
        }
        __syncthreads();  // every read of this round's wdst and sbase is done before the next walk
    }
}

}  // namespace

hipError_t launch_exchange_count(const XchgLaunch& a, hipStream_t s) {
    if (a.ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL(xchg_count_kernel, dim3(a.ntiles), dim3(kXThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_exchange_scan(const XchgLaunch& a, hipStream_t s) {
    hipLaunchKernelGGL(xchg_scan_kernel, dim3(a.shares), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_exchange_write(const XchgLaunch& a, hipStream_t s) {
    if (a.ntiles == 0) return hipSuccess;
    hipLaunchKernelGGL(xchg_write_kernel, dim3(a.ntiles), dim3(kXThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace ofp
