// ofp_cell.hip -- the per-cell frontier transfer step for gfx950.
//
// One cell p of the Grid-FP sweep (reference semantics:
// src/common/gridfp_transition.hpp:113-146, src/common/gridfp_reverse.hpp:39-98,
// applied as in tests/factor_transfer_batch_test.cu:60-72) is performed IN
// PLACE on the whole device-resident state with one thread per blocked word.
//
// Why one thread per blocked word: write a blocked word of width W-1 at cell p
// as b = [P, x, S] (P = positions above p-1, x = position p-1, S = positions
// below).  The main words touched by cell p pair up with b exactly:
//   x = N : main [P,N,N,S] (NN) and [P,L,R,S] (LR);  NN --include--> LR
//   x = R : main [P,N,R,S] (NR) and [P,R,N,S] (RN);  RN --include--> NR
//   x = L : main [P,N,L,S] (NL) and [P,L,N,S] (LN);  LN --include--> NL
// and the blocked input consumed at cell p (blocked_exclude = minsert(b,p,N))
// lands on [P,N,x,S], the blocked output produced at cell p has the same index
// b.  Every main word written at cell p (NN, LR, NR, NL; at p=1 also RN) is
// owned by exactly one thread, and every main word read by a foreign thread
// (RL, LL, RR: sources of closures) is never written at cell p.  Hence one
// launch updates main and blocked in place with no atomics and no second copy
// of the state; at n=22 with 14 moduli that is 156 GB instead of 313 GB.
//
// Index math: the thread decodes b by walking the height table from the top
// (lexicographic rank, N<R<L, entered at height 1), accumulating along the way
// the rank contribution the prefix P has in the main numbering.  Every main
// index it needs differs from that base by the contributions of the two cell
// positions (and, for closure predecessors, by one flipped label), so each is
// a few table lookups.  The closure predecessors of an x=N target are the
// reference's reverse_block_predecessors (LL: top-level L's of S; RR:
// top-level R's of P), enumerated with running rank sums so that each costs
// O(1) beyond the scan.
//
// Residues: uint32 mod p, additions only (every transfer coefficient is 1);
// a + b mod p is one conditional subtract, written as a subtract with borrow
// and a select (addm in ofp_cell_impl.h).
#include <cstring>
#include <vector>

#include "ofp_cell_impl.h"

namespace ofp {

namespace {

using namespace cell;

// Copy the first kLdsRows x C entries of a kDpRows x kDpCols table into LDS
// (row stride C there).
template <int C, typename T>
__device__ void load_table(T* dst, const T* src) {
    for (int t = threadIdx.x; t < kLdsRows * C; t += blockDim.x) dst[t] = src[(t / C) * kDpCols + t % C];
}

// Work distribution: the blocked range is cut into chunks of blockDim.x words;
// workgroup g runs on XCD g%8 (round-robin dispatch).  The chunks are dealt to
// the XCDs in pieces of 2^shift consecutive chunks, round-robin (piece i to XCD
// i%8), and XCD x's workgroups stride through its pieces in order; shift < 0
// gives each XCD one contiguous eighth.  Each XCD's L2 sees contiguous pieces,
// and work whose cost depends on where in the rank order it lies (at the first
// cells the x=N words, which scan for closures, are the first third of the
// blocked range) is shared by all XCDs.
template <int KC, typename Idx, int L, int V>
__global__ __launch_bounds__(256) void cell_kernel(CellLaunch a) {
    __shared__ Quad<Idx> tab[kLdsRows][LdsCols<Idx>::value];
    const Quad<Idx>* g = (const Quad<Idx>*)a.dp;
    load_table<LdsCols<Idx>::value>(&tab[0][0], g);
    __syncthreads();
    Tab<Idx> T{tab};
    const Direct<KC, L, Idx> mem{a.main, a.blk, a.K, a.off};
    const Sched sc(a.nblk, blockDim.x, gridDim.x, blockIdx.x, a.xcd_shift);
    for (uint64_t t = sc.slot; t < sc.count; t += sc.nx) {
        const uint64_t b = sc.chunk(t) * blockDim.x + threadIdx.x;
        if (b < a.nblk) cell_one<KC, Idx, L, V>(a, T, a.first + b, mem);
    }
}

// Sorted assignment (state-major, uniform orbit body): the 256 words of a chunk
// are decoded, then dealt back to the workgroup's threads with the x=N words
// first, so the closure scans (x=N only) fill whole waves and every wave loads
// and stores one kind of orbit word for consecutive words.  Each word still has
// exactly one owner, so the in-place argument is unchanged.  Three workgroup
// barriers per chunk; slot and cnt are the workgroup's LDS.
template <int KC, typename Idx, class Mem>
__device__ void sorted_sweep(const CellLaunch& a, const Tab<Idx>& T, const Mem& mem, Decoded<Idx>* slot,
                             uint32_t (*cnt)[4]) {
    const Sched sc(a.nblk, blockDim.x, gridDim.x, blockIdx.x, a.xcd_shift);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;  // lanes under this one
    for (uint64_t t = sc.slot; t < sc.count; t += sc.nx) {
        const uint64_t b = sc.chunk(t) * blockDim.x + threadIdx.x;
        Decoded<Idx> d{};
        int key = 2;  // 0: x = N, 1: x = R or L, 2: past the end
        if (b < a.nblk) {
            d = decode_blocked<Idx>(T, a.W, a.p, a.first + b);
            key = d.x == 0 ? 0 : 1;
        }
        const uint64_t m0 = __ballot(key == 0), m1 = __ballot(key == 1);
        if (lane == 0) {
            cnt[0][w] = (uint32_t)__popcll(m0);
            cnt[1][w] = (uint32_t)__popcll(m1);
        }
        __syncthreads();
        uint32_t o0 = 0, o1 = 0, n0 = 0, n1 = 0;
        for (uint32_t i = 0; i < nw; ++i) {
            if (i < w) {
                o0 += cnt[0][i];
                o1 += cnt[1][i];
            }
            n0 += cnt[0][i];
            n1 += cnt[1][i];
        }
        if (key == 0) slot[o0 + (uint32_t)__popcll(m0 & below)] = d;
        else if (key == 1) slot[n0 + o1 + (uint32_t)__popcll(m1 & below)] = d;
        __syncthreads();
        if (threadIdx.x < n0 + n1) {
            const Decoded<Idx> e = slot[threadIdx.x];
            cell_body<KC, Idx, kAoS, 2>(a, T, e, mem);
        }
        __syncthreads();  // slot and cnt are rewritten for the next chunk
    }
}

template <int KC, typename Idx>
__global__ __launch_bounds__(256) void cell_sorted_kernel(CellLaunch a) {
    __shared__ Quad<Idx> tab[kLdsRows][LdsCols<Idx>::value];
    __shared__ Decoded<Idx> slot[256];
    __shared__ uint32_t cnt[2][4];
    const Quad<Idx>* g = (const Quad<Idx>*)a.dp;
    load_table<LdsCols<Idx>::value>(&tab[0][0], g);
    __syncthreads();
    Tab<Idx> T{tab};
    const Direct<KC, kAoS, Idx> mem{a.main, a.blk, a.K, a.off};
    sorted_sweep<KC, Idx>(a, T, mem, slot, cnt);
}

// Shard variant: same per-thread body, windowed memory policy (state-major).
template <int KC, typename Idx, int V>
__global__ __launch_bounds__(256) void cell_shard_kernel(CellLaunch a) {
    __shared__ Quad<Idx> tab[kLdsRows][LdsCols<Idx>::value];
    const Quad<Idx>* g = (const Quad<Idx>*)a.dp;
    load_table<LdsCols<Idx>::value>(&tab[0][0], g);
    __syncthreads();
    Tab<Idx> T{tab};
    Windowed<KC> mem;
    mem.main = a.main;
    mem.blk = a.blk;
    mem.m0 = a.m0;
    mem.M = a.M;
    mem.b0 = a.first;
    mem.seed = a.seed;
    mem.K = a.K;
    mem.off = a.off;
#pragma unroll
    for (int j = 0; j < KC; ++j) mem.mod[j] = a.mod[j];
    mem.stray = a.stray;
    const Sched sc(a.nblk, blockDim.x, gridDim.x, blockIdx.x, a.xcd_shift);
    for (uint64_t t = sc.slot; t < sc.count; t += sc.nx) {
        const uint64_t b = sc.chunk(t) * blockDim.x + threadIdx.x;
        if (b < a.nblk) cell_one<KC, Idx, kAoS, V>(a, T, a.first + b, mem);
    }
}

// Capped variant: the threads enumerate, in rank order, only the blocked words
// whose heights stay <= the launch's cap (their count is a.nblk), decoded with
// the capped completion table; every other blocked word is left untouched.
// The uniform orbit body (V = 2) runs on each decoded word.
template <int KC, typename Idx>
__global__ __launch_bounds__(256) void cell_capped_kernel(CellLaunch a) {
    __shared__ Quad<Idx> tab[kLdsRows][LdsCols<Idx>::value];
    __shared__ Pair<Idx> ctab[kLdsRows][LdsCols<Idx>::value];
    const Quad<Idx>* g = (const Quad<Idx>*)a.dp;
    const Pair<Idx>* gc = (const Pair<Idx>*)a.dpc;
    load_table<LdsCols<Idx>::value>(&tab[0][0], g);
    load_table<LdsCols<Idx>::value>(&ctab[0][0], gc);
    __syncthreads();
    Tab<Idx> T{tab};
    const Direct<KC, kAoS, Idx> mem{a.main, a.blk, a.K, a.off};
    const Sched sc(a.nblk, blockDim.x, gridDim.x, blockIdx.x, a.xcd_shift);
    for (uint64_t u = sc.slot; u < sc.count; u += sc.nx) {
        const uint64_t t = sc.chunk(u) * blockDim.x + threadIdx.x;
        if (t < a.nblk) cell_body<KC, Idx, kAoS, 2>(a, T, decode_capped<Idx>(T, ctab, a.W, a.p, t), mem);
    }
}

__global__ void fill_remote_kernel(uint32_t* main, uint64_t m0, uint64_t M, uint32_t* blk, uint64_t b0,
                                   uint64_t B, int K, int k, const uint32_t* mods, uint64_t seed) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (M + B) * (uint64_t)k; t += stride) {
        const uint64_t s = t / (uint64_t)k;
        const int j = (int)(t - s * (uint64_t)k);
        if (s < M) main[s * K + j] = remote_residue(seed, m0 + s, j, mods[j]);
        else blk[(s - M) * K + j] = remote_residue(seed, kBlockedDomain + b0 + (s - M), j, mods[j]);
    }
}

__global__ void set_one_kernel(uint32_t* main, uint64_t index, int K, int k, int layout,
                               const uint32_t* mods) {
    int j = threadIdx.x;
    if (j < k) main[state_elem(layout, index, K, j)] = 1u % mods[j];
}

template <typename Idx, int KC>
hipError_t launch_kc(const CellLaunch& a, int threads, int blocks, hipStream_t s) {
    if (a.layout == kAoS) {
        if (a.variant == 1)
            hipLaunchKernelGGL((cell_kernel<KC, Idx, kAoS, 1>), dim3(blocks), dim3(threads), 0, s, a);
        else if (a.variant == 2)
            hipLaunchKernelGGL((cell_kernel<KC, Idx, kAoS, 2>), dim3(blocks), dim3(threads), 0, s, a);
        else if (a.variant == 3)
            hipLaunchKernelGGL((cell_kernel<KC, Idx, kAoS, 3>), dim3(blocks), dim3(threads), 0, s, a);
        else if (a.variant == 4)
            hipLaunchKernelGGL((cell_sorted_kernel<KC, Idx>), dim3(blocks), dim3(threads), 0, s, a);
        else
            hipLaunchKernelGGL((cell_kernel<KC, Idx, kAoS, 0>), dim3(blocks), dim3(threads), 0, s, a);
    } else {
        hipLaunchKernelGGL((cell_kernel<KC, Idx, kAoSoA, 0>), dim3(blocks), dim3(threads), 0, s, a);
    }
    return hipGetLastError();
}

template <typename Idx, int KC>
hipError_t launch_shard_kc(const CellLaunch& a, int threads, int blocks, hipStream_t s) {
    if (a.variant == 1)
        hipLaunchKernelGGL((cell_shard_kernel<KC, Idx, 1>), dim3(blocks), dim3(threads), 0, s, a);
    else if (a.variant == 2 || a.variant == 4)  // shards are never sorted: the uniform body
        hipLaunchKernelGGL((cell_shard_kernel<KC, Idx, 2>), dim3(blocks), dim3(threads), 0, s, a);
    else if (a.variant == 3)
        hipLaunchKernelGGL((cell_shard_kernel<KC, Idx, 3>), dim3(blocks), dim3(threads), 0, s, a);

    else
        hipLaunchKernelGGL((cell_shard_kernel<KC, Idx, 0>), dim3(blocks), dim3(threads), 0, s, a);
    return hipGetLastError();
}

template <typename Idx>
hipError_t launch_shard_idx(const CellLaunch& a, int threads, int blocks, hipStream_t s) {
    switch (a.kc) {
#define OFP_KC(n) case n: return launch_shard_kc<Idx, n>(a, threads, blocks, s);
        OFP_KC(1) OFP_KC(2) OFP_KC(4) OFP_KC(8)
#undef OFP_KC
    default: return hipErrorInvalidValue;
    }
}

template <typename Idx>
hipError_t launch_capped_idx(const CellLaunch& a, int threads, int blocks, hipStream_t s) {
    switch (a.kc) {
#define OFP_KC(n) case n: hipLaunchKernelGGL((cell_capped_kernel<n, Idx>), dim3(blocks), dim3(threads), 0, s, a); break;
        OFP_KC(1) OFP_KC(2) OFP_KC(3) OFP_KC(4) OFP_KC(5) OFP_KC(6) OFP_KC(7) OFP_KC(8)
        OFP_KC(9) OFP_KC(10) OFP_KC(11) OFP_KC(12) OFP_KC(13) OFP_KC(14) OFP_KC(15) OFP_KC(16)
#undef OFP_KC
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename Idx>
hipError_t launch_idx(const CellLaunch& a, int threads, int blocks, hipStream_t s) {
    switch (a.kc) {
#define OFP_KC(n) case n: return launch_kc<Idx, n>(a, threads, blocks, s);
        OFP_KC(1) OFP_KC(2) OFP_KC(3) OFP_KC(4) OFP_KC(5) OFP_KC(6) OFP_KC(7) OFP_KC(8)
        OFP_KC(9) OFP_KC(10) OFP_KC(11) OFP_KC(12) OFP_KC(13) OFP_KC(14) OFP_KC(15) OFP_KC(16)
#undef OFP_KC
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

hipError_t launch_cell(const CellLaunch& a, bool wide_index, int threads, int blocks,
                       hipStream_t s) {
    if (threads <= 0 || threads > 256 || (threads & 63) || blocks <= 0 || (blocks & 7))
        return hipErrorInvalidConfiguration;
    if (a.W < 2 || a.W > kMaxW || a.p < 1 || a.p > a.W - 1 || a.kc < 1 || a.kc > kMaxKC ||
        a.off < 0 || a.off + a.kc > a.K || (a.layout != kAoSoA && a.layout != kAoS) ||
        (a.layout == kAoS && a.K > 1 && ((a.K & 1) || (a.off & 1))))
        return hipErrorInvalidValue;
    return wide_index ? launch_idx<uint64_t>(a, threads, blocks, s)
                      : launch_idx<uint32_t>(a, threads, blocks, s);
}

hipError_t launch_cell_capped(const CellLaunch& a, bool wide_index, int threads, int blocks,
                              hipStream_t s) {
    if (threads <= 0 || threads > 256 || (threads & 63) || blocks <= 0 || (blocks & 7))
        return hipErrorInvalidConfiguration;
    if (a.W < 2 || a.W > kMaxW || a.p < 1 || a.p > a.W - 1 || a.kc < 1 || a.kc > kMaxKC || a.off < 0 ||
        a.off + a.kc > a.K || a.layout != kAoS || (a.K > 1 && ((a.K & 1) || (a.off & 1))) || !a.dpc ||
        a.first != 0)
        return hipErrorInvalidValue;
    return wide_index ? launch_capped_idx<uint64_t>(a, threads, blocks, s)
                      : launch_capped_idx<uint32_t>(a, threads, blocks, s);
}

hipError_t launch_cell_shard(const CellLaunch& a, bool wide_index, int threads, int blocks,
                             hipStream_t s) {
    if (threads <= 0 || threads > 256 || (threads & 63) || blocks <= 0 || (blocks & 7))
        return hipErrorInvalidConfiguration;
    if (a.W < 2 || a.W > kMaxW || a.p < 1 || a.p > a.W - 1 || a.layout != kAoS || a.off < 0 ||
        a.off + a.kc > a.K || (a.K > 1 && ((a.K & 1) || (a.off & 1))) || !a.stray)
        return hipErrorInvalidValue;
    return wide_index ? launch_shard_idx<uint64_t>(a, threads, blocks, s)
                      : launch_shard_idx<uint32_t>(a, threads, blocks, s);
}

hipError_t launch_fill_remote(uint32_t* main, uint64_t m0, uint64_t M, uint32_t* blk, uint64_t b0,
                              uint64_t B, int K, int k, const uint32_t* mods_dev, uint64_t seed,
                              hipStream_t s) {
    if (k < 1 || K < k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fill_remote_kernel, dim3(4096), dim3(256), 0, s, main, m0, M, blk, b0, B, K, k,
                       mods_dev, seed);
    return hipGetLastError();
}

struct ModVals {
    uint32_t m[kMaxKC];
};
__global__ void set_one_values_kernel(uint32_t* main, uint64_t index, int K, int k, int layout, ModVals mods) {
    int j = threadIdx.x;
    if (j < k) main[state_elem(layout, index, K, j)] = 1u % mods.m[j];
}

hipError_t launch_set_one_values(uint32_t* main, uint64_t index, int K, int k, int layout,
                                 const uint32_t* mods_host, hipStream_t s) {
    if (k < 1 || k > kMaxKC || K < k) return hipErrorInvalidValue;
    ModVals mv{};
    for (int j = 0; j < k; ++j) mv.m[j] = mods_host[j];
    hipLaunchKernelGGL(set_one_values_kernel, dim3(1), dim3(64), 0, s, main, index, K, k, layout, mv);
    return hipGetLastError();
}

hipError_t launch_set_one(uint32_t* main, uint64_t index, int K, int k, int layout,
                          const uint32_t* mods_dev, hipStream_t s) {
    if (k < 1 || k > 1024 || K < k) return hipErrorInvalidValue;
    hipLaunchKernelGGL(set_one_kernel, dim3(1), dim3(((k + 63) / 64) * 64), 0, s, main, index, K, k,
                       layout, mods_dev);
    return hipGetLastError();
}

// Packed into a properly aligned buffer, then copied out byte-wise (out may be
// any host buffer).
template <typename Idx>
static size_t pack_into(const uint64_t (*d)[kDpCols], void* out, size_t cap) {
    const size_t bytes = sizeof(Quad<Idx>) * kDpRows * kDpCols;
    if (cap < bytes) return 0;
    std::vector<Quad<Idx>> q(kDpRows * kDpCols);
    pack_quads<Idx>(d, (Quad<Idx>(*)[kDpCols])q.data());
    std::memcpy(out, q.data(), bytes);
    return bytes;
}

size_t pack_height_table(bool wide, void* out, size_t cap) {
    uint64_t d[kDpRows][kDpCols];
    build_height_table(d);
    return wide ? pack_into<uint64_t>(d, out, cap) : pack_into<uint32_t>(d, out, cap);
}

size_t pack_group_table(bool high, int W, int low, uint32_t mask, void* out, size_t cap) {
    const size_t bytes = sizeof(Quad<uint32_t>) * kDpRows * kDpCols;
    if (cap < bytes) return 0;
    Occupancy om, ob;
    group_occupancy(high, W, low, mask, om, ob);
    uint64_t fm[kDpRows][kDpCols], fb[kDpRows][kDpCols];
    build_group_table(om, fm);
    build_group_table(ob, fb);
    std::vector<Quad<uint32_t>> q(kDpRows * kDpCols);
    pack_group_quads<uint32_t>(ob, fb, om, fm, (Quad<uint32_t>(*)[kDpCols])q.data());
    std::memcpy(out, q.data(), bytes);
    return bytes;
}

void group_counts(bool high, int W, int low, uint32_t mask, uint64_t* main_words, uint64_t* blocked_words) {
    Occupancy om, ob;
    group_occupancy(high, W, low, mask, om, ob);
    uint64_t f[kDpRows][kDpCols];
    build_group_table(om, f);
    *main_words = f[W][1];
    build_group_table(ob, f);
    *blocked_words = f[W - 1][1];
}

void build_capped_table(int cap, uint64_t (*d)[kDpCols]) {
    for (int w = 0; w < kDpRows; ++w)
        for (int h = 0; h < kDpCols; ++h) d[w][h] = 0;
    d[0][0] = 1;
    for (int w = 1; w < kDpRows; ++w)
        for (int h = 0; h <= cap && h < kDpCols; ++h) {
            uint64_t x = d[w - 1][h];
            if (h > 0) x += d[w - 1][h - 1];
            if (h + 1 <= cap && h + 1 < kDpCols) x += d[w - 1][h + 1];
            d[w][h] = x;
        }
}

size_t pack_capped_tables(bool wide, void* out, size_t cap) {
    const size_t one = (wide ? sizeof(Pair<uint64_t>) : sizeof(Pair<uint32_t>)) * kDpRows * kDpCols;
    if (cap < one * kMaxCap) return 0;
    uint64_t d[kDpRows][kDpCols];
    std::vector<Pair<uint64_t>> q64(kDpRows * kDpCols);
    std::vector<Pair<uint32_t>> q32(kDpRows * kDpCols);
    for (int c = 1; c <= kMaxCap; ++c) {
        build_capped_table(c, d);
        char* dst = (char*)out + one * (c - 1);
        if (wide) {
            pack_pairs<uint64_t>(d, (Pair<uint64_t>(*)[kDpCols])q64.data());
            std::memcpy(dst, q64.data(), one);
        } else {
            pack_pairs<uint32_t>(d, (Pair<uint32_t>(*)[kDpCols])q32.data());
            std::memcpy(dst, q32.data(), one);
        }
    }
    return one * kMaxCap;
}

void build_height_table(uint64_t (*d)[kDpCols]) {
    for (int w = 0; w < kDpRows; ++w)
        for (int h = 0; h < kDpCols; ++h) d[w][h] = 0;
    d[0][0] = 1;
    for (int w = 1; w < kDpRows; ++w)
        for (int h = 0; h < kDpCols; ++h) {
            uint64_t x = d[w - 1][h];
            if (h > 0) x += d[w - 1][h - 1];
            if (h + 1 < kDpCols) x += d[w - 1][h + 1];
            d[w][h] = x;
        }
}

}  // namespace ofp
