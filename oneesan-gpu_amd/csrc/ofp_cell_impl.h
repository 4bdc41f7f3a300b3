// ofp_cell_impl.h -- per-thread body of the cell step (see ofp_cell.hip for
// the design).  Kept in a header with host+device qualifiers so that the same
// index math can be exercised by a host harness.
#pragma once
#include "ofp_internal.h"

#ifndef OFP_HD
#define OFP_HD __host__ __device__ __forceinline__
#endif

namespace ofp {
namespace cell {

OFP_HD uint32_t addm(uint32_t a, uint32_t b, uint32_t m) {
    uint32_t s = a + b;
    return (s < a || s >= m) ? s - m : s;
}

template <typename Idx>
struct Tab {
    const Idx (*d)[kDpCols];
    OFP_HD Idx at(int pos, int h) const { return d[pos][h]; }
    // rank contribution of value v at position pos entered at height h
    OFP_HD Idx c(int pos, int v, int h) const {
        Idx z = 0;
        if (v >= 1) z = d[pos][h];
        if (v == 2 && h > 0) z += d[pos][h - 1];
        return z;
    }
    OFP_HD Idx cL(int pos, int h) const {
        return d[pos][h] + (h > 0 ? d[pos][h - 1] : Idx(0));
    }
    // decode one position: consume r, return value (0=N,1=R,2=L) and its contribution
    OFP_HD int step(int pos, int h, Idx& r, Idx& contrib) const {
        Idx zN = d[pos][h];
        if (r < zN) { contrib = 0; return 0; }
        r -= zN;
        Idx zR = h > 0 ? d[pos][h - 1] : Idx(0);
        if (r < zR) { contrib = zN; return 1; }
        r -= zR;
        contrib = zN + zR;
        return 2;
    }
};

template <int KC>
struct Vals {
    uint32_t v[KC];
};

template <int KC>
OFP_HD void load_vals(const uint32_t* __restrict__ base, uint64_t i, int K,
                                          int off, Vals<KC>& out) {
    const uint32_t* q = base + aosoa(i, K, off);
#pragma unroll
    for (int j = 0; j < KC; ++j) out.v[j] = q[j * kTile];
}
template <int KC>
OFP_HD void store_vals(uint32_t* __restrict__ base, uint64_t i, int K, int off,
                                           const Vals<KC>& in) {
    uint32_t* q = base + aosoa(i, K, off);
#pragma unroll
    for (int j = 0; j < KC; ++j) q[j * kTile] = in.v[j];
}
template <int KC>
OFP_HD void acc_vals(Vals<KC>& acc, const uint32_t* __restrict__ base,
                                         uint64_t i, int K, int off, const uint32_t* mod) {
    const uint32_t* q = base + aosoa(i, K, off);
#pragma unroll
    for (int j = 0; j < KC; ++j) acc.v[j] = addm(acc.v[j], q[j * kTile], mod[j]);
}

template <int KC, typename Idx>
OFP_HD void cell_one(const CellLaunch& a, const Tab<Idx>& T, uint64_t bidx) {
    const int W = a.W, p = a.p, K = a.K, off = a.off;
    uint32_t* __restrict__ mainp = a.main;
    uint32_t* __restrict__ blkp = a.blk;
    uint32_t mod[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) mod[j] = a.mod[j];

    // ---- decode blocked word b over positions W-2..p-1 ----
    Idx r = (Idx)bidx;
    int h = 1;
    Idx mainbase = 0;  // sum of main-numbering contributions of P
    uint64_t word = 0;
    for (int pos = W - 2; pos >= p; --pos) {
        Idx cb;
        int v = T.step(pos, h, r, cb);
        mainbase += T.c(pos + 1, v, h);
        word |= (uint64_t)v << (2 * pos);
        h += (v == 2) - (v == 1);
    }
    const int hP = h;
    Idx cx;
    const int x = T.step(p - 1, hP, r, cx);
    const Idx sr = r;  // rank of S among suffixes entered at the height after x

    if (x == 0) {
        // orbit NN -> LR
        const uint64_t iNN = (uint64_t)(mainbase + sr);
        const uint64_t iLR = (uint64_t)(mainbase + T.cL(p, hP) + T.at(p - 1, hP + 1) + sr);
        Vals<KC> nn, lr, bc, acc;
        load_vals<KC>(mainp, iNN, K, off, nn);
        load_vals<KC>(mainp, iLR, K, off, lr);
        if (a.read_blk) load_vals<KC>(blkp, bidx, K, off, bc);
#pragma unroll
        for (int j = 0; j < KC; ++j) acc.v[j] = 0;

        if (p >= 2) {
            // RL predecessor: [P,R,L,S]
            if (hP > 0) {
                const uint64_t iRL = (uint64_t)(mainbase + T.at(p, hP) + T.cL(p - 1, hP - 1) + sr);
                acc_vals<KC>(acc, mainp, iRL, K, off, mod);
            }
            // LL predecessors: [P,L,L,S*], S* = S with a top-level L at q turned into R,
            // entered at height hP+2 (reverse_block_predecessors, gridfp_reverse.hpp:62-68)
            if (p >= 3) {
                const Idx baseLL = mainbase + T.cL(p, hP) + T.cL(p - 1, hP + 1);
                Idx rr = sr, s0 = 0, s2 = 0;
                int hh = hP, depth = 1;
                for (int pos = p - 2; pos >= 0; --pos) {
                    Idx c0;
                    int v = T.step(pos, hh, rr, c0);
                    if (v == 2 && depth == 1) {
                        const uint64_t iLL = (uint64_t)(baseLL + s2 + T.at(pos, hh + 2) + (sr - s0 - c0));
                        acc_vals<KC>(acc, mainp, iLL, K, off, mod);
                    }
                    s0 += c0;
                    s2 += T.c(pos, v, hh + 2);
                    if (v == 2) ++depth;
                    else if (v == 1) { if (--depth == 0) break; }
                    hh += (v == 2) - (v == 1);
                }
            }
        }
        // RR predecessors: [P*,R,R,S], P* = P with a top-level R at q turned into L
        // (gridfp_reverse.hpp:70-82); at p=1 they are the closure sources of NN.
        {
            const Idx tail = T.at(p, hP + 2) + T.at(p - 1, hP + 1) + sr;
            Idx u0 = 0, u2 = 0;
            int depth = 1, Ha = hP;
            for (int mp = p + 1; mp <= W - 1; ++mp) {
                const int v = (int)((word >> (2 * (mp - 1))) & 3u);
                const int H = Ha + (v == 1) - (v == 2);
                const Idx c0 = T.c(mp, v, H);
                if (v == 1 && depth == 1) {
                    const uint64_t iRR = (uint64_t)(mainbase - (u0 + c0) + T.cL(mp, H) + u2 + tail);
                    acc_vals<KC>(acc, mainp, iRR, K, off, mod);
                }
                u0 += c0;
                u2 += T.c(mp, v, H + 2);
                if (v == 2) { if (--depth == 0) break; }
                else if (v == 1) ++depth;
                Ha = H;
            }
        }
        Vals<KC> onn, olr;
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            uint32_t t = a.read_blk ? addm(nn.v[j], bc.v[j], mod[j]) : nn.v[j];
            olr.v[j] = addm(lr.v[j], nn.v[j], mod[j]);
            if (p == 1) t = addm(t, acc.v[j], mod[j]);
            onn.v[j] = t;
        }
        store_vals<KC>(mainp, iNN, K, off, onn);
        store_vals<KC>(mainp, iLR, K, off, olr);
        if (p >= 2) store_vals<KC>(blkp, bidx, K, off, acc);
    } else if (x == 1) {
        // orbit RN -> NR (and at p=1 also NR -> RN)
        const uint64_t iNR = (uint64_t)(mainbase + T.at(p - 1, hP) + sr);
        const uint64_t iRN = (uint64_t)(mainbase + T.at(p, hP) + sr);
        Vals<KC> nr, rn, bc, o;
        load_vals<KC>(mainp, iNR, K, off, nr);
        load_vals<KC>(mainp, iRN, K, off, rn);
        if (a.read_blk) load_vals<KC>(blkp, bidx, K, off, bc);
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            uint32_t t = addm(nr.v[j], rn.v[j], mod[j]);
            if (a.read_blk) t = addm(t, bc.v[j], mod[j]);
            o.v[j] = t;
        }
        store_vals<KC>(mainp, iNR, K, off, o);
        if (p >= 2) {
            store_vals<KC>(blkp, bidx, K, off, nr);
        } else {
#pragma unroll
            for (int j = 0; j < KC; ++j) o.v[j] = addm(rn.v[j], nr.v[j], mod[j]);
            store_vals<KC>(mainp, iRN, K, off, o);
        }
    } else {
        // orbit LN -> NL (cannot occur at p=1)
        const uint64_t iNL = (uint64_t)(mainbase + T.cL(p - 1, hP) + sr);
        const uint64_t iLN = (uint64_t)(mainbase + T.cL(p, hP) + sr);
        Vals<KC> nl, ln, bc, o;
        load_vals<KC>(mainp, iNL, K, off, nl);
        load_vals<KC>(mainp, iLN, K, off, ln);
        if (a.read_blk) load_vals<KC>(blkp, bidx, K, off, bc);
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            uint32_t t = addm(nl.v[j], ln.v[j], mod[j]);
            if (a.read_blk) t = addm(t, bc.v[j], mod[j]);
            o.v[j] = t;
        }
        store_vals<KC>(mainp, iNL, K, off, o);
        if (p >= 2) store_vals<KC>(blkp, bidx, K, off, nl);
    }
}

}  // namespace cell
}  // namespace ofp
