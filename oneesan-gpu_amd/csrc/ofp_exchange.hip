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
// at most kXTileWords words of one group, each round of a wave taking 64
// consecutive local ranks.  Words with one top part (the fixed HIGH positions)
// are a run of consecutive ranks -- a block, about a hundred words at W = 28 --
// with one destination share and one top-part HIGH-rank contribution, so a
// wave decodes top parts only in rounds that leave its current block
// (xchg::decode_top) and takes each word's bottom contribution from the
// group's bottom table (xchg::build_bottom_lut):
//   count  -- count each word's destination share (one wave64 ballot per share
//             and round);
//   scan   -- one workgroup per share: exclusive prefix of the tile counts;
//   write  -- write each word's record.  Records keep the window order inside a
//             bucket: each round of 256 words is compacted with one wave64
//             ballot per share (a lane's slot = the tile's cursor + the set
//             lanes below it + the counts of the lower waves), so the buckets
//             are deterministic.
// The residues are read once in window order (coalesced) and written once as
// records; records of one share in one round are contiguous.
#include "ofp_exchange_impl.h"

namespace ofp {

namespace {

using cell::Quad;
constexpr int kXThreads = 256;
constexpr int kXWaves = kXThreads / 64;
constexpr int kXSlices = 4;  // slices of 256 words per round of the write pass

__device__ void load_quads(Quad<uint32_t> (*dst)[kDpCols], const Quad<uint32_t>* src) {
    Quad<uint32_t>* d = &dst[0][0];
    for (int t = threadIdx.x; t < kDpRows * kDpCols; t += blockDim.x) d[t] = src[t];
}

__device__ const Quad<uint32_t>* group_table(const XchgLaunch& a, uint32_t g) {
    return reinterpret_cast<const Quad<uint32_t>*>(a.gtab + (size_t)g * sizeof(Quad<uint32_t>) * kDpRows * kDpCols);
}

// A block: the consecutive local ranks whose words share one top part (the
// fixed HIGH positions), hence one HIGH mask, destination share and top-part
// HIGH-rank contribution; a word's bottom contributes lut[lut + r - start].
struct Blk {
    uint32_t start, len, mask, top, lut;
    int share;
};

// Rank is false in the count pass, which needs a block's extent and share only
// (not its top-part HIGH rank, whose group-row reads are global loads).
template <bool Main, bool Rank>
__device__ Blk find_block(const XchgLaunch& a, const Quad<uint32_t> (*tq)[kDpCols], const uint32_t* off, uint32_t r) {
    const xchg::Top t = xchg::decode_top<Main>(tq, a.top, a.base, r);
    Blk b;
    b.start = r - t.r_rest;
    b.lut = off[t.h];
    b.len = off[t.h + 1] - b.lut;
    b.mask = t.mask;
    b.top = Rank ? xchg::top_high(t, a.ht) : 0u;
    b.share = xchg::share_of(t.mask, a.first, a.shares);
    return b;
}

// Give every valid lane of the wave the block of its rank r.  `cb` is the
// wave's current block (the same in every lane), kept from round to round:
// rounds walk consecutive ranks, so in a large block every lane falls in it and
// nothing is decoded.  Otherwise the lanes outside it decode their own top parts
// side by side (one decode's time; blocks of a few words are common where the
// group's occupied positions are few), and the block of the last valid lane,
// which the next round continues, becomes the wave's current block.
template <bool Rank>
__device__ void assign_blocks(const XchgLaunch& a, const Quad<uint32_t> (*tq)[kDpCols], const uint32_t* off,
                              bool valid, uint32_t r, Blk& cb, Blk& mine) {
    const bool in = valid && (r - cb.start < cb.len);
    mine = cb;
    if (__ballot(valid && !in) == 0) return;
    if (valid && !in) mine = a.main ? find_block<true, Rank>(a, tq, off, r) : find_block<false, Rank>(a, tq, off, r);
    const uint64_t vm = __ballot(valid);
    const int last = 63 - __clzll((long long)vm);
    cb.start = (uint32_t)__shfl((int)mine.start, last);
    cb.len = (uint32_t)__shfl((int)mine.len, last);
    cb.mask = (uint32_t)__shfl((int)mine.mask, last);
    if (Rank) cb.top = (uint32_t)__shfl((int)mine.top, last);
    cb.lut = (uint32_t)__shfl((int)mine.lut, last);
    cb.share = __shfl(mine.share, last);
}

// The group's bottom offsets (kDpCols + 1 entries) into LDS.
__device__ void load_off(uint32_t* dst, const XchgLaunch& a, uint32_t g) {
    const uint32_t* src = a.lut_off + (size_t)g * (kDpCols + 1);
    for (int t = threadIdx.x; t < kDpCols + 1; t += blockDim.x) dst[t] = src[t];
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
    // lane q of each wave counts share q
    const int lane = threadIdx.x & 63;
    uint32_t mine_cnt = 0;
    Blk cb{0u, 0u, 0u, 0u, 0u, 0};
    for (uint32_t r0 = 0; r0 < t.len; r0 += kXThreads) {
        const uint32_t i = r0 + threadIdx.x;
        const bool valid = i < t.len;
        Blk mine;
        assign_blocks<false>(a, tq, off, valid, t.rank0 + i, cb, mine);
        const int s = valid ? mine.share : -1;
        const uint64_t vm = __ballot(valid);
        if (vm == 0) continue;
        // the usual case: one destination share for the whole round
        const int s0 = __shfl(s, __ffsll((long long)vm) - 1);
        if (__ballot(valid && s != s0) == 0) {
            if (lane == s0) mine_cnt += (uint32_t)__popcll(vm);
            continue;
        }
        for (int q = 0; q < a.shares; ++q) {
            const uint32_t c = (uint32_t)__popcll(__ballot(s == q));
            if (lane == q) mine_cnt += c;
        }
    }
    if (lane < a.shares) atomicAdd(&cnt[lane], mine_cnt);
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

// Rounds of kXSlices x 256 words: slice j of a round is the 256 consecutive
// ranks r0 + 256 j ..; each thread loads its words' residues for every slice
// first (so that a round keeps kXSlices loads per thread in flight), walks
// the blocks and counts the slices' records per destination share and wave,
// then one thread per share turns the counts of the round's (slice, wave)
// pairs, in window order, into absolute bucket slots, and the words are stored.
__global__ __launch_bounds__(kXThreads) void xchg_write_kernel(XchgLaunch a) {
    constexpr int E = kXSlices * kXWaves;  // (slice, wave) pairs of a round, in window order
    __shared__ Quad<uint32_t> tq[kDpRows][kDpCols];
    __shared__ uint64_t cursor[kMaxShares];
    __shared__ uint32_t wcnt[E][kMaxShares];
    __shared__ uint64_t sbase[E][kMaxShares];
    __shared__ uint32_t off[kDpCols + 1];
    const XTile t = a.tiles[blockIdx.x];
    load_quads(tq, group_table(a, t.group));
    load_off(off, a, t.group);
    Blk cb{0u, 0u, 0u, 0u, 0u, 0};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < a.shares) {
        uint64_t start = 0;
        for (int q = 0; q < tid; ++q) start += a.totals[q];
        cursor[tid] = start + a.offs[(size_t)tid * a.ntiles + blockIdx.x];
    }
    __syncthreads();  // the tables and cursors are in LDS before any wave decodes
    const uint64_t below = (1ull << lane) - 1ull;
    const bool even = (a.K & 1) == 0;
    for (uint32_t r0 = 0; r0 < t.len; r0 += kXThreads * kXSlices) {
        uint2 v01[kXSlices], v23[kXSlices];
        int sh[kXSlices];
        uint32_t pre[kXSlices];
        uint64_t dst[kXSlices];
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
#pragma unroll
        for (int j = 0; j < kXSlices; ++j) {
            const uint32_t i = r0 + (uint32_t)(j * kXThreads + tid);
            const bool valid = i < t.len;
            const uint32_t r = t.rank0 + i;
            Blk mine;
            assign_blocks<true>(a, tq, off, valid, r, cb, mine);
            sh[j] = -1;
            dst[j] = 0;
            if (valid) {
                sh[j] = mine.share;
                dst[j] = a.base_of[mine.mask] + (uint32_t)(mine.top + a.lut[mine.lut + (r - mine.start)]);
            }
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
        // (the next round's first barrier orders these reads of sbase before
        // its rewrite)
#pragma unroll
        for (int j = 0; j < kXSlices; ++j) {
            if (sh[j] < 0) continue;
            const uint32_t i = r0 + (uint32_t)(j * kXThreads + tid);
            const uint64_t slot = sbase[j * kXWaves + wave][sh[j]] + pre[j];
            a.out_idx[slot] = dst[j];
            const uint32_t* src = a.res + (t.word0 + i) * (uint64_t)a.K;
            uint32_t* out = a.out_res + slot * (uint64_t)a.K;
            if (even) {
                *reinterpret_cast<uint2*>(out) = v01[j];
                if (a.K >= 4) *reinterpret_cast<uint2*>(out + 2) = v23[j];
                for (int k = 4; k < a.K; k += 2)
                    *reinterpret_cast<uint2*>(out + k) = *reinterpret_cast<const uint2*>(src + k);
            } else {
                for (int k = 0; k < a.K; ++k) out[k] = src[k];
            }
        }
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
