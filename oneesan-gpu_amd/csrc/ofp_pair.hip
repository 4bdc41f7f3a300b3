// ofp_pair.hip -- the pair pass: cells p and p-1 of a row applied in place,
// one launch per window height change d (body and design in ofp_pair_impl.h).
//
// Each thread takes one class (P', d, S'), decoded from its class number with
// the class completion table of (p, d) and the height table, both in LDS; the
// classes are dealt to the XCDs in pieces exactly as the cell kernels deal
// their blocked words (cell::Sched).  Every word belongs to one class and only
// that class's thread writes it; the words a class reads from other classes
// still hold their values from before the pass when it reads them (launch
// order d = -1, 0, 1, then -2, 2), so the launches need no atomics and no
// second copy of the state.
#include <cstring>

#include "ofp_pair_impl.h"

namespace ofp {

namespace {

using namespace cell;

template <int C, typename T>
__device__ void load_rows(T* dst, const T* src) {
    for (int t = threadIdx.x; t < kLdsRows * C; t += blockDim.x) dst[t] = src[t];
}

// MB: workgroups of 256 per CU the register allocation must allow (2: at most
// 256 VGPRs, no spills at 14 moduli; 3: at most 168, with spills).
template <int KC, typename Idx, int D, int GF, int MB>
__global__ __launch_bounds__(256, MB) void pair_kernel(pair::PairLaunch a) {
    constexpr int C = LdsCols<Idx>::value;
    __shared__ Quad<Idx> tab[kLdsRows][C];
    __shared__ Pair<Idx> ctab[kLdsRows][C];
    const Quad<Idx>* g = (const Quad<Idx>*)a.dp;
    for (int t = threadIdx.x; t < kLdsRows * C; t += blockDim.x) (&tab[0][0])[t] = g[(t / C) * kDpCols + t % C];
    load_rows<C>(&ctab[0][0], (const Pair<Idx>*)a.ct);
    __syncthreads();
    Tab<Idx> T{tab};
    const pair::Io<KC, Idx> io{a.main_in, a.blk_in, a.main_out, a.blk_out, a.K, a.off};
    const Sched sc(a.classes, blockDim.x, gridDim.x, blockIdx.x, a.xcd_shift);
    for (uint64_t t = sc.slot; t < sc.count; t += sc.nx) {
        const uint64_t c = sc.chunk(t) * blockDim.x + threadIdx.x;
        if (c < a.classes) {
            const pair::Class<Idx> k = pair::decode_class<Idx>(T, ctab, a.W, a.p, c);
            pair::pair_body<KC, Idx, D, GF>(a, T, k, io);
        }
    }
}

template <int KC, typename Idx, int GF, int MB>
hipError_t launch_d(const pair::PairLaunch& a, int threads, int blocks, hipStream_t s) {
    switch (a.d) {
        case -3: hipLaunchKernelGGL((pair_kernel<KC, Idx, -3, GF, MB>), dim3(blocks), dim3(threads), 0, s, a); break;
        case -2: hipLaunchKernelGGL((pair_kernel<KC, Idx, -2, GF, MB>), dim3(blocks), dim3(threads), 0, s, a); break;
        case -1: hipLaunchKernelGGL((pair_kernel<KC, Idx, -1, GF, MB>), dim3(blocks), dim3(threads), 0, s, a); break;
        case 0: hipLaunchKernelGGL((pair_kernel<KC, Idx, 0, GF, MB>), dim3(blocks), dim3(threads), 0, s, a); break;
        case 1: hipLaunchKernelGGL((pair_kernel<KC, Idx, 1, GF, MB>), dim3(blocks), dim3(threads), 0, s, a); break;
        case 2: hipLaunchKernelGGL((pair_kernel<KC, Idx, 2, GF, MB>), dim3(blocks), dim3(threads), 0, s, a); break;
        case 3: hipLaunchKernelGGL((pair_kernel<KC, Idx, 3, GF, MB>), dim3(blocks), dim3(threads), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Pair-kernel variant (PairLaunch::variant): bit 0 = gathers first, bit 1 = the
// three-workgroups-per-CU register budget; only the 14- and 16-moduli kernels
// carry the variants (the others run variant 0).
template <int KC, typename Idx>
hipError_t launch_gf(const pair::PairLaunch& a, int threads, int blocks, hipStream_t s) {
    if constexpr (KC >= 14) {
        switch (a.variant & 3) {
            case 1: return launch_d<KC, Idx, 1, 2>(a, threads, blocks, s);
            case 2: return launch_d<KC, Idx, 0, 3>(a, threads, blocks, s);
            case 3: return launch_d<KC, Idx, 1, 3>(a, threads, blocks, s);
            default: break;
        }
    }
    return launch_d<KC, Idx, 0, 2>(a, threads, blocks, s);
}

}  // namespace

size_t pack_class_tables(bool wide, void* out, size_t cap) {
    uint64_t d0[kDpRows][kDpCols];
    build_height_table(d0);
    const size_t one = (wide ? sizeof(Pair<uint64_t>) : sizeof(Pair<uint32_t>)) * kLdsRows *
                       (wide ? LdsCols<uint64_t>::value : LdsCols<uint32_t>::value);
    const size_t bytes = one * kMaxW * 7;
    if (cap < bytes) return 0;
    char* o = static_cast<char*>(out);
    for (int p = 0; p < kMaxW; ++p)
        for (int d = -3; d <= 3; ++d) {
            char* dst = o + one * (size_t)(p * 7 + d + 3);
            if (wide) {
                static Pair<uint64_t> t[kDpRows][LdsCols<uint64_t>::value];
                pair::pack_class_table<uint64_t, LdsCols<uint64_t>::value>(d0, p, d, t);
                std::memcpy(dst, t, one);
            } else {
                static Pair<uint32_t> t[kDpRows][LdsCols<uint32_t>::value];
                pair::pack_class_table<uint32_t, LdsCols<uint32_t>::value>(d0, p, d, t);
                std::memcpy(dst, t, one);
            }
        }
    return bytes;
}

size_t class_table_bytes(bool wide) {
    return (wide ? sizeof(Pair<uint64_t>) * LdsCols<uint64_t>::value : sizeof(Pair<uint32_t>) * LdsCols<uint32_t>::value) *
           kLdsRows;
}

uint64_t pair_class_count(int W, int p, int d) {
    static uint64_t d0[kDpRows][kDpCols];
    static bool ready = false;
    if (!ready) {
        build_height_table(d0);
        ready = true;
    }
    return pair::class_count(d0, W, p, d);
}

hipError_t launch_pair(const pair::PairLaunch& a, bool wide, int threads, int blocks, hipStream_t s) {
    if (a.p < 2 || a.p > a.W - 1 || a.d < -3 || a.d > 3 || a.kc < 1 || a.off + a.kc > a.K) return hipErrorInvalidValue;
    if (wide) {
        switch (a.kc) {
            case 1: return launch_gf<1, uint64_t>(a, threads, blocks, s);
            case 2: return launch_gf<2, uint64_t>(a, threads, blocks, s);
            case 4: return launch_gf<4, uint64_t>(a, threads, blocks, s);
            case 6: return launch_gf<6, uint64_t>(a, threads, blocks, s);
            case 8: return launch_gf<8, uint64_t>(a, threads, blocks, s);
            case 10: return launch_gf<10, uint64_t>(a, threads, blocks, s);
            case 12: return launch_gf<12, uint64_t>(a, threads, blocks, s);
            case 14: return launch_gf<14, uint64_t>(a, threads, blocks, s);
            case 16: return launch_gf<16, uint64_t>(a, threads, blocks, s);
            default: return hipErrorInvalidValue;
        }
    }
    switch (a.kc) {
        case 1: return launch_gf<1, uint32_t>(a, threads, blocks, s);
        case 2: return launch_gf<2, uint32_t>(a, threads, blocks, s);
        case 4: return launch_gf<4, uint32_t>(a, threads, blocks, s);
        case 6: return launch_gf<6, uint32_t>(a, threads, blocks, s);
        case 8: return launch_gf<8, uint32_t>(a, threads, blocks, s);
        case 10: return launch_gf<10, uint32_t>(a, threads, blocks, s);
        case 12: return launch_gf<12, uint32_t>(a, threads, blocks, s);
        case 14: return launch_gf<14, uint32_t>(a, threads, blocks, s);
        case 16: return launch_gf<16, uint32_t>(a, threads, blocks, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace ofp
