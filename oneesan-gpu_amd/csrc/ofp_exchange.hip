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
// at most kXTileWords words of one group:
//   count  -- decode the fixed positions of each word and count its
//             destination share (one wave64 ballot per share and round);
//   scan   -- one workgroup per share: exclusive prefix of the tile counts;
//   write  -- decode each word whole, ranking it in its HIGH group on the way
//             (xchg::decode_rank), and write its record.  Records keep the
//             window order inside a bucket: each round of 256 words is
//             compacted with one wave64 ballot per share (a lane's slot = the
//             tile's cursor + the set lanes below it + the counts of the lower
//             waves), so the buckets are deterministic.
// The residues are read once in window order (coalesced) and written once as
// records; records of one share in one round are contiguous.
#include "ofp_exchange_impl.h"

namespace ofp {

namespace {

using cell::Quad;
constexpr int kXThreads = 256;
constexpr int kXWaves = kXThreads / 64;

__device__ void load_quads(Quad<uint32_t> (*dst)[kDpCols], const Quad<uint32_t>* src) {
    Quad<uint32_t>* d = &dst[0][0];
    for (int t = threadIdx.x; t < kDpRows * kDpCols; t += blockDim.x) d[t] = src[t];
}

__device__ const Quad<uint32_t>* group_table(const XchgLaunch& a, uint32_t g) {
    return reinterpret_cast<const Quad<uint32_t>*>(a.gtab + (size_t)g * sizeof(Quad<uint32_t>) * kDpRows * kDpCols);
}

__global__ __launch_bounds__(kXThreads) void xchg_count_kernel(XchgLaunch a) {
    __shared__ Quad<uint32_t> tq[kDpRows][kDpCols];
    __shared__ uint32_t cnt[kMaxShares];
    const XTile t = a.tiles[blockIdx.x];
    load_quads(tq, group_table(a, t.group));
    if (threadIdx.x < kMaxShares) cnt[threadIdx.x] = 0;
    __syncthreads();
    // lane q of each wave counts share q
    const int lane = threadIdx.x & 63;
    uint32_t mine = 0;
    for (uint32_t r0 = 0; r0 < t.len; r0 += kXThreads) {
        const uint32_t i = r0 + threadIdx.x;
        int s = -1;
        if (i < t.len) {
            const uint64_t w = a.main ? xchg::decode_group_word<true>(tq, a.top, a.base, t.rank0 + i)
                                      : xchg::decode_group_word<false>(tq, a.top, a.base, t.rank0 + i);
            s = xchg::share_of(xchg::high_mask(w, a.base, a.nf), a.first, a.shares);
        }
        for (int q = 0; q < a.shares; ++q) {
            const uint32_t c = (uint32_t)__popcll(__ballot(s == q));
            if (lane == q) mine += c;
        }
    }
    if (lane < a.shares) atomicAdd(&cnt[lane], mine);
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

__global__ __launch_bounds__(kXThreads) void xchg_write_kernel(XchgLaunch a) {
    __shared__ Quad<uint32_t> tq[kDpRows][kDpCols];
    __shared__ Quad<uint32_t> dq[kDpRows][kDpCols];
    __shared__ uint64_t cursor[kMaxShares];
    __shared__ uint32_t wcnt[kXWaves][kMaxShares];
    const XTile t = a.tiles[blockIdx.x];
    load_quads(tq, group_table(a, t.group));
    load_quads(dq, reinterpret_cast<const Quad<uint32_t>*>(a.dplain));
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid < a.shares) {
        uint64_t start = 0;
        for (int q = 0; q < tid; ++q) start += a.totals[q];
        cursor[tid] = start + a.offs[(size_t)tid * a.ntiles + blockIdx.x];
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t r0 = 0; r0 < t.len; r0 += kXThreads) {
        const uint32_t i = r0 + tid;
        const bool valid = i < t.len;
        int s = -1;
        uint64_t dst = 0;
        if (valid) {
            uint64_t w;
            uint32_t mask;
            const uint32_t rank = a.main ? xchg::decode_rank<true>(tq, dq, a.ht, a.top, a.base, t.rank0 + i, w, mask)
                                         : xchg::decode_rank<false>(tq, dq, a.ht, a.top, a.base, t.rank0 + i, w, mask);
            s = xchg::share_of(mask, a.first, a.shares);
            dst = a.base_of[mask] + rank;
        }
        uint32_t pre = 0;
        for (int q = 0; q < a.shares; ++q) {
            const uint64_t m = __ballot(s == q);
            if (s == q) pre = (uint32_t)__popcll(m & below);
            if (lane == 0) wcnt[wave][q] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (valid) {
            uint64_t slot = cursor[s] + pre;
            for (int v = 0; v < wave; ++v) slot += wcnt[v][s];
            a.out_idx[slot] = dst;
            const uint32_t* src = a.res + (t.word0 + i) * (uint64_t)a.K;
            uint32_t* out = a.out_res + slot * (uint64_t)a.K;
            if ((a.K & 1) == 0) {
                for (int j = 0; j < a.K; j += 2)
                    *reinterpret_cast<uint2*>(out + j) = *reinterpret_cast<const uint2*>(src + j);
            } else {
                for (int j = 0; j < a.K; ++j) out[j] = src[j];
            }
        }
        __syncthreads();
        if (tid < a.shares) {
            uint32_t add = 0;
            for (int v = 0; v < kXWaves; ++v) add += wcnt[v][tid];
            cursor[tid] += add;
        }
        __syncthreads();
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
