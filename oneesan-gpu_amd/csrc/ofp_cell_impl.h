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

// a + b mod m for a, b < m <= 2^32 - 1, without forming the possibly wrapping
// sum: a + b >= m exactly when a >= m - b, and then a - (m - b) is the answer.
// The borrow of a - (m - b) is that comparison, so the whole is four VALU
// operations (m - b from a scalar, subtract with borrow, add, select).
OFP_HD uint32_t addm(uint32_t a, uint32_t b, uint32_t m) {
    const uint32_t mb = m - b;
    uint32_t t;
    const bool lt = __builtin_sub_overflow(a, mb, &t);
    return lt ? a + b : t;
}

// One entry of the packed height table, for position pos and height h, with
// D[.][-1] = 0:
//   zn = D[pos][h], zc = D[pos][h] + D[pos][h-1]            -- decode at pos
//   mn = D[pos+1][h], mc = D[pos+1][h] + D[pos+1][h-1]      -- the same one position up
// zn is the rank contribution of an R at pos (the N completions it passes), zc
// that of an L (N and R completions), so a walk step compares against and adds
// table values only.  Decoding a blocked word (width W-1) and accumulating its
// main-numbering rank (width W, every prefix position shifted up by one) needs
// one 16-byte (or 32-byte) table read per position.  With 32-bit Idx the sums
// are taken mod 2^32, exactly as a sum of the two truncated counts would be.
template <typename Idx>
struct alignas(4 * sizeof(Idx)) Quad {
    Idx zn, zc, mn, mc;
};

// Whole-entry table reads (16 or 32 bytes at once).  Left to itself the
// compiler reads only the fields each branch of a step needs, one dependent
// LDS access after another; the decode walks and closure scans are chains of
// such steps, so one wide read per step shortens every chain.
template <typename Idx>
OFP_HD Quad<Idx> ldq(const Quad<Idx>& e) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(Idx) == 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(&e);
        return Quad<Idx>{v.x, v.y, v.z, v.w};
    } else {
        const ulonglong2* q = reinterpret_cast<const ulonglong2*>(&e);
        const ulonglong2 a = q[0], b = q[1];
        return Quad<Idx>{(Idx)a.x, (Idx)a.y, (Idx)b.x, (Idx)b.y};
    }
#else
    return e;
#endif
}
// The first half only (zn, zc): what the closure scans need.
template <typename Idx>
struct Z2 {
    Idx zn, zc;
};
template <typename Idx>
OFP_HD Z2<Idx> ldz(const Quad<Idx>& e) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(Idx) == 4) {
        const uint2 v = *reinterpret_cast<const uint2*>(&e);
        return Z2<Idx>{v.x, v.y};
    } else {
        const ulonglong2 v = *reinterpret_cast<const ulonglong2*>(&e);
        return Z2<Idx>{(Idx)v.x, (Idx)v.y};
    }
#else
    return Z2<Idx>{e.zn, e.zc};
#endif
}
// The second half only (mn, mc): the entries of the main position one up.
template <typename Idx>
OFP_HD Z2<Idx> ldm(const Quad<Idx>& e) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(Idx) == 4) {
        const uint2 v = *(reinterpret_cast<const uint2*>(&e) + 1);
        return Z2<Idx>{v.x, v.y};
    } else {
        const ulonglong2 v = *(reinterpret_cast<const ulonglong2*>(&e) + 1);
        return Z2<Idx>{(Idx)v.x, (Idx)v.y};
    }
#else
    return Z2<Idx>{e.mn, e.mc};
#endif
}
// Rank contribution of value v (0=N, 1=R, 2=L) from a table entry's pair.
template <typename Idx>
OFP_HD Idx pick(int v, Idx n, Idx c) {
    return v == 0 ? Idx(0) : (v == 1 ? n : c);
}
// One position of a rank walk without branches: v = number of completion
// blocks (N, then R) that r passes; contrib = their size; r -= contrib.
// zc = zn + D[pos][h-1] never wraps: it is at most D[pos+1][h], a word count
// of the state.
template <typename Idx>
OFP_HD int step_z(Idx zn, Idx zc, Idx& r, Idx& contrib) {
    const bool g0 = r >= zn, g1 = r >= zc;
    contrib = g1 ? zc : (g0 ? zn : Idx(0));
    r -= contrib;
    return (int)g0 + (int)g1;
}

// Rows and columns of a height table kept in LDS.  A word of W <= kMaxW positions
// entered at height 1 is at height <= 14 at every position it can still be
// completed from, and the scans read at most two heights above that, so columns
// 0..21 and positions 0..kMaxW-1 hold every entry read.  At 64-bit
// ranks that keeps the quad table at 19.7 KB instead of 30 KB, so eight workgroups
// of 256 fit in a CU's 160 KB of LDS; 32-bit tables keep all columns.
constexpr int kLdsRows = kMaxW;
template <typename Idx>
struct LdsCols {
    static constexpr int value = sizeof(Idx) == 8 ? 22 : kDpCols;
};

template <typename Idx>
struct Tab {
    const Quad<Idx> (*q)[LdsCols<Idx>::value];
    OFP_HD Idx at(int pos, int h) const { return q[pos][h].zn; }
    OFP_HD Idx cL(int pos, int h) const { return q[pos][h].zc; }
    // The same for main position pos >= 1, from the main half of row pos-1.
    // A cell p reads the entries of main positions >= p this way: the rows
    // themselves hold blocked-numbering entries, which at those positions count
    // the blocked words (one position shorter) of an occupancy group whose
    // constrained positions lie above the cell.
    OFP_HD Idx mat(int pos, int h) const { return q[pos - 1][h].mn; }
    OFP_HD Idx mcL(int pos, int h) const { return q[pos - 1][h].mc; }
    // rank contribution of value v (0=N,1=R,2=L) at position pos entered at height h
    OFP_HD Idx c(int pos, int v, int h) const {
        const Quad<Idx> e = q[pos][h];
        return pick(v, e.zn, e.zc);
    }
    // decode one position: consume r, return the value and its contribution
    OFP_HD int step(int pos, int h, Idx& r, Idx& contrib) const {
        const Z2<Idx> e = ldz(q[pos][h]);
        return step_z(e.zn, e.zc, r, contrib);
    }
    OFP_HD static int step_e(const Quad<Idx>& e, Idx& r, Idx& contrib) {
        return step_z(e.zn, e.zc, r, contrib);
    }
};

// Chunk schedule of one workgroup over n words in chunks of bdim (see the
// work-distribution note in ofp_cell.hip).
struct Sched {
    uint64_t count;  // chunks of this workgroup's XCD (multiple of the piece size)
    uint64_t chunks;
    uint32_t xcd, slot, nx;
    int shift;
    OFP_HD Sched(uint64_t n, uint32_t bdim, uint32_t nwg, uint32_t block, int sh)
        : chunks((n + bdim - 1) / bdim), xcd(block & 7u), slot(block >> 3), nx(nwg >> 3), shift(sh) {
        if (shift < 0) {
            count = ((chunks * (xcd + 1)) >> 3) - ((chunks * xcd) >> 3);
        } else {
            const uint64_t pieces = (chunks + (1ull << shift) - 1) >> shift;
            count = (pieces > xcd ? (pieces - xcd + 7) >> 3 : 0) << shift;
        }
    }
    // the t-th chunk of this XCD; may be >= chunks in a last, partial piece
    OFP_HD uint64_t chunk(uint64_t t) const {
        if (shift < 0) return ((chunks * xcd) >> 3) + t;
        return ((((t >> shift) << 3) + xcd) << shift) + (t & ((1ull << shift) - 1));
    }
};

// State layouts.  kAoSoA: tiles of 64 states, modulus-major inside a tile
// (a wave reading one modulus of 64 consecutive states reads 256 contiguous
// bytes).  kAoS: state-major with an even stride, residues of one state
// contiguous (one scattered state is one or two 64-byte sectors), read two
// residues per 8-byte load.
enum Layout : int { kAoSoA = 0, kAoS = 1 };

template <int L>
OFP_HD uint64_t elem(uint64_t i, int K, int j) {
    if constexpr (L == kAoSoA) return aosoa(i, K, j);
    else return i * (uint64_t)(uint32_t)K + (uint64_t)(uint32_t)j;
}

template <int KC>
struct Vals {
    uint32_t v[KC];
};

template <int KC, int L>
OFP_HD void load_vals(const uint32_t* __restrict__ base, uint64_t i, int K, int off, Vals<KC>& out) {
    const uint32_t* q = base + elem<L>(i, K, off);
    if constexpr (L == kAoSoA) {
#pragma unroll
        for (int j = 0; j < KC; ++j) out.v[j] = q[j * kTile];
    } else {
        const uint2* q2 = reinterpret_cast<const uint2*>(q);
#pragma unroll
        for (int j = 0; j + 1 < KC; j += 2) {
            uint2 t = q2[j / 2];
            out.v[j] = t.x;
            out.v[j + 1] = t.y;
        }
        if constexpr (KC & 1) out.v[KC - 1] = q[KC - 1];
    }
}
template <int KC, int L>
OFP_HD void store_vals(uint32_t* __restrict__ base, uint64_t i, int K, int off, const Vals<KC>& in) {
    uint32_t* q = base + elem<L>(i, K, off);
    if constexpr (L == kAoSoA) {
#pragma unroll
        for (int j = 0; j < KC; ++j) q[j * kTile] = in.v[j];
    } else {
        uint2* q2 = reinterpret_cast<uint2*>(q);
#pragma unroll
        for (int j = 0; j + 1 < KC; j += 2) q2[j / 2] = make_uint2(in.v[j], in.v[j + 1]);
        if constexpr (KC & 1) q[KC - 1] = in.v[KC - 1];
    }
}
template <int KC>
OFP_HD void add_into(Vals<KC>& acc, const Vals<KC>& x, const uint32_t* mod) {
#pragma unroll
    for (int j = 0; j < KC; ++j) acc.v[j] = addm(acc.v[j], x.v[j], mod[j]);
}

// Memory policy of a launch: Direct addresses the whole device-resident state.
// Idx: the index width of the launch; a 32-bit index makes every element
// offset one 32x32->64 multiply-add.
template <int KC, int L, typename Idx = uint64_t>
struct Direct {
    uint32_t* __restrict__ main;
    uint32_t* __restrict__ blk;
    int K, off;
    OFP_HD void lm(uint64_t i, Vals<KC>& v) const { load_vals<KC, L>(main, (uint64_t)(Idx)i, K, off, v); }
    OFP_HD void sm(uint64_t i, const Vals<KC>& v) const { store_vals<KC, L>(main, (uint64_t)(Idx)i, K, off, v); }
    OFP_HD void lb(uint64_t b, Vals<KC>& v) const { load_vals<KC, L>(blk, (uint64_t)(Idx)b, K, off, v); }
    OFP_HD void sb(uint64_t b, const Vals<KC>& v) const { store_vals<KC, L>(blk, (uint64_t)(Idx)b, K, off, v); }
};

// Synthetic residues of states that live on other shards: a seeded hash of the
// global index (blocked words in their own domain), mapped onto [0, p) by a
// multiply-high.  32-bit arithmetic only, so that standing in for a read costs
// little more than the read would.  The CPU oracle evaluates the same function.
OFP_HD uint32_t remote_residue(uint64_t seed, uint64_t index, int j, uint32_t mod) {
    uint32_t h = (uint32_t)index * 0x9E3779B1u ^ ((uint32_t)(index >> 32) + (uint32_t)seed) * 0x85EBCA77u ^
                 ((uint32_t)j + (uint32_t)(seed >> 32)) * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (uint32_t)(((uint64_t)h * mod) >> 32);
}
constexpr uint64_t kBlockedDomain = 1ull << 62;

// Memory policy of a shard launch (state-major layout): main words in the
// resident window [m0, m0+M) and blocked words in [b0, b0+B) are in HBM; main
// words outside the window belong to other shards and read as remote_residue().
// A shard made of whole prefix classes owns every word a cell below its window
// boundary writes, so stores outside the window do not occur; they are counted
// (an invariant the tests check) and dropped.
template <int KC>
struct Windowed {
    uint32_t* __restrict__ main;
    uint32_t* __restrict__ blk;
    uint64_t m0, M, b0;
    uint64_t seed;
    int K, off;
    uint32_t mod[KC];
    unsigned long long* stray;  // count of out-of-window stores (must stay 0)
    OFP_HD void lm(uint64_t i, Vals<KC>& v) const {
        if (i - m0 < M) load_vals<KC, 1>(main, i - m0, K, off, v);
        else {
#pragma unroll
            for (int j = 0; j < KC; ++j) v.v[j] = remote_residue(seed, i, off + j, mod[j]);
        }
    }
    OFP_HD void sm(uint64_t i, const Vals<KC>& v) const {
        if (i - m0 < M) store_vals<KC, 1>(main, i - m0, K, off, v);
        else {
#ifdef __HIP_DEVICE_COMPILE__
            atomicAdd(stray, 1ull);
#else
            ++*stray;
#endif
        }
    }
    OFP_HD void lb(uint64_t b, Vals<KC>& v) const { load_vals<KC, 1>(blk, b - b0, K, off, v); }
    OFP_HD void sb(uint64_t b, const Vals<KC>& v) const { store_vals<KC, 1>(blk, b - b0, K, off, v); }
};

// Closure predecessors are found one at a time by a scan; their residues are
// accumulated one candidate late: candidate c's values are added when candidate
// c+1 is found (or at finish), so the scan runs on while c's loads are in flight.
// The loads go straight into `pend` after its previous contents are added: a
// load into a temporary followed by a copy into `pend` would make the copy wait
// for the load at once.
template <int KC, int L>
struct Gather {
    Vals<KC> acc, pend;
    bool have = false;
    OFP_HD void init() {
#pragma unroll
        for (int j = 0; j < KC; ++j) acc.v[j] = 0;
        have = false;
    }
    template <class Mem>
    OFP_HD void add(const Mem& m, uint64_t i, const uint32_t* mod) {
        if (have) add_into<KC>(acc, pend, mod);
        m.lm(i, pend);
        have = true;
    }
    template <class Mem>
    OFP_HD void finish(const Mem&, const uint32_t* mod) {
        if (have) add_into<KC>(acc, pend, mod);
        have = false;
    }
};

// Closure predecessors of the blocked word [P, N, S] (x = N): RL, LL and RR
// sources, accumulated into g (the reverse of gridfp_transition.hpp:113-146's
// closure cases; order and multiplicity as gridfp_reverse.hpp:39-98).
template <int KC, typename Idx, int L, class Mem, class G>
OFP_HD void gather_closures(const Tab<Idx>& T, int W, int p, int hP, Idx mainbase, Idx sr, uint64_t word,
                            const Mem& m, G& g, const uint32_t* mod) {
    if (p >= 2) {
        // RL predecessor: [P,R,L,S]
        if (hP > 0) {
            const uint64_t iRL = (uint64_t)(mainbase + T.mat(p, hP) + T.cL(p - 1, hP - 1) + sr);
            g.add(m, iRL, mod);
        }
        // LL predecessors: [P,L,L,S*], S* = S with a top-level L at q turned into R,
        // entered at height hP+2 (reverse_block_predecessors, gridfp_reverse.hpp:62-68)
        // The scan walks S downwards decoding it from its rank rr; an L opens and
        // an R closes one level, so the bracket depth below x is 1 + hh - hP and
        // the top level is hh == hP.  After a step rr is the rank of the rest of
        // S, which S* keeps; s2 is the rank of S*'s part above q (entered two
        // levels higher).
        if (p >= 3) {
            const Idx baseLL = mainbase + T.mcL(p, hP) + T.cL(p - 1, hP + 1);
            Idx rr = sr, s2 = 0;
            int hh = hP;
            for (int pos = p - 2; pos >= 0; --pos) {
                const Z2<Idx> e = ldz(T.q[pos][hh]);
                const Z2<Idx> e2 = ldz(T.q[pos][hh + 2]);
                const bool g0 = rr >= e.zn, g1 = rr >= e.zc;  // L: g1; R: g0 and not g1
                rr -= g1 ? e.zc : (g0 ? e.zn : Idx(0));
                if (g1 && hh == hP) g.add(m, (uint64_t)(baseLL + s2 + e2.zn + rr), mod);
                if (g0 && !g1 && hh == hP) break;
                s2 += g1 ? e2.zc : (g0 ? e2.zn : Idx(0));
                hh += g1 ? 1 : (g0 ? -1 : 0);
            }
        }
    }
    // RR predecessors: [P*,R,R,S], P* = P with a top-level R at q turned into L
    // (gridfp_reverse.hpp:70-82); at p=1 they are the closure sources of NN.
    // The scan walks P upwards from its lowest position; an R opens and an L
    // closes one level, so the depth is 1 + Ha - hP and the top level Ha == hP.
    // ub is the main-numbering base without the part of P scanned so far, u2 the
    // rank of that part in P* (entered two levels higher below q).
    {
        Idx ub = mainbase + T.mat(p, hP + 2) + T.at(p - 1, hP + 1) + sr;
        Idx u2 = 0;
        int Ha = hP;
        uint64_t w = word >> (2 * p);  // P from its lowest position up
        for (int mp = p + 1; mp <= W - 1; ++mp) {
            const uint32_t v = (uint32_t)w & 3u;
            w >>= 2;
            const bool isR = v == 1, isL = v == 2;
            const int H = Ha + (isR ? 1 : (isL ? -1 : 0));
            const Z2<Idx> e0 = ldm(T.q[mp - 1][H]);  // main position mp
            const Z2<Idx> e2 = ldm(T.q[mp - 1][H + 2]);
            ub -= isL ? e0.zc : (isR ? e0.zn : Idx(0));
            if (isR && Ha == hP) g.add(m, (uint64_t)(ub + e0.zc + u2), mod);
            if (isL && Ha == hP) break;
            u2 += isL ? e2.zc : (isR ? e2.zn : Idx(0));
            Ha = H;
        }
    }
}

// V selects when an x=N thread loads its own orbit values (NN, LR, blocked-in):
// 0 = before the closure scans (more loads in flight, more registers live during
// the scans), 1 = after them (fewer live registers, so more waves per SIMD),
// 2 = one branch-free orbit path for all three kinds of blocked word, loads first,
// 3 = as 2 but the orbit results are stored before the closure scans.
// A blocked word b = [P, x, S] of cell p, decoded: its index, the height hP
// entering x, x itself, the main-numbering rank contribution of P (mainbase),
// the rank of S among suffixes entered at the height after x (sr) and the word
// (2 bits per blocked position; the closure scans read P from it).
template <typename Idx>
struct Decoded {
    uint64_t bidx;
    Idx mainbase, sr;
    uint64_t word;
    int hP, x;
};

// Decode blocked word bidx over positions W-2..p-1 (S is left as its rank).
template <typename Idx>
OFP_HD Decoded<Idx> decode_blocked(const Tab<Idx>& T, int W, int p, uint64_t bidx) {
    Decoded<Idx> d;
    d.bidx = bidx;
    Idx r = (Idx)bidx;
    int h = 1;
    Idx mainbase = 0;
    uint64_t word = 0;  // built top-down, then moved to bits 2*pos
    for (int pos = W - 2; pos >= p; --pos) {
        const Quad<Idx> e = ldq(T.q[pos][h]);
        // the two comparisons select everything: v = g0 + g1, and R (g0 only)
        // lowers the height by one, L (both) raises it by one
        const bool g0 = r >= e.zn, g1 = r >= e.zc;
        r -= g1 ? e.zc : (g0 ? e.zn : Idx(0));
        mainbase += g1 ? e.mc : (g0 ? e.mn : Idx(0));
        word = (word << 2) | (uint64_t)((int)g0 + (int)g1);
        h += g1 ? 1 : (g0 ? -1 : 0);
    }
    d.hP = h;
    Idx cx;
    d.x = T.step(p - 1, h, r, cx);
    d.sr = r;
    d.mainbase = mainbase;
    d.word = word << (2 * p);
    return d;
}

// Capped completions for the words whose heights stay <= a cap c: at position
// pos entered at height h, n = Dc[pos][h] (after choosing N), c = Dc[pos][h] +
// Dc[pos][h-1] (after choosing N or R), Dc[.][-1] = 0.
template <typename Idx>
struct alignas(2 * sizeof(Idx)) Pair {
    Idx n, c;
};

// Decode the t-th (in rank order) blocked word whose heights stay <= the cap
// of C.  Symbols are chosen with the capped counts; every rank (the blocked
// index, P's main-numbering contribution, S's rank) is summed with the full
// table T, so the result is the same as decode_blocked(T, W, p, bidx) for the
// word's own index bidx.
template <typename Idx>
OFP_HD Decoded<Idx> decode_capped(const Tab<Idx>& T, const Pair<Idx> (*C)[LdsCols<Idx>::value], int W, int p,
                                   uint64_t t) {
    Decoded<Idx> d;
    Idx r = (Idx)t, bidx = 0, mainbase = 0, sr = 0;
    uint64_t word = 0;
    int h = 1, x = 0, hP = 1;
    for (int pos = W - 2; pos >= 0; --pos) {
        const Pair<Idx> pc = C[pos][h];
        const Quad<Idx> e = ldq(T.q[pos][h]);
        Idx skip;
        const int v = step_z(pc.n, pc.c, r, skip);
        const Idx cb = pick(v, e.zn, e.zc);
        bidx += cb;
        if (pos >= p) mainbase += pick(v, e.mn, e.mc);
        else if (pos == p - 1) { x = v; hP = h; }
        else sr += cb;
        word |= (uint64_t)v << (2 * pos);
        h += (v == 2) - (v == 1);
    }
    d.bidx = (uint64_t)bidx;
    d.mainbase = mainbase;
    d.sr = sr;
    d.word = word;
    d.hP = hP;
    d.x = x;
    return d;
}

template <int KC, typename Idx, int L, int V, class Mem>
OFP_HD void cell_body(const CellLaunch& a, const Tab<Idx>& T, const Decoded<Idx>& d, const Mem& m) {
    const int W = a.W, p = a.p;
    uint32_t mod[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) mod[j] = a.mod[j];
    const uint64_t bidx = d.bidx;
    const int hP = d.hP, x = d.x;
    const Idx mainbase = d.mainbase, sr = d.sr;
    const uint64_t word = d.word;

    if constexpr (V >= 2) {
        // One code path for the three orbits: every lane issues its orbit loads
        // (A, B and blocked-in) before the x = N lanes scan for closure sources,
        // so at small p, where a wave holds all three kinds of blocked word, the
        // loads of R and L lanes overlap the scans instead of following them.
        //   x = N: A = NN, B = LR;  A' = NN + blk (+ closures at p = 1), B' = LR + NN
        //   x = R: A = NR, B = RN;  A' = NR + RN + blk, at p = 1 B' = RN + NR
        //   x = L: A = NL, B = LN;  A' = NL + LN + blk
        // blocked-out (p >= 2) = closure sum for x = N, old A otherwise.
        const Idx cLp = T.mcL(p, hP);
        Idx offA, offB;
        if (x == 0) {
            offA = 0;
            offB = cLp + T.at(p - 1, hP + 1);
        } else if (x == 1) {
            offA = T.at(p - 1, hP);
            offB = T.mat(p, hP);
        } else {
            offA = T.cL(p - 1, hP);
            offB = cLp;
        }
        const uint64_t iA = (uint64_t)(mainbase + offA + sr);
        const uint64_t iB = (uint64_t)(mainbase + offB + sr);
        Vals<KC> va, vb, bc;
        m.lm(iA, va);
        m.lm(iB, vb);
        if (a.read_blk) m.lb(bidx, bc);
        if constexpr (V == 3) {
            // Early retire: the orbit words and the R/L lanes' blocked-out are
            // stored before the closure scans, so only the closure sum (and, at
            // p = 1, NN's partial sum) stays live across them.
            Vals<KC> oa, ob;
#pragma unroll
            for (int j = 0; j < KC; ++j) {
                uint32_t t = va.v[j];
                if (a.read_blk) t = addm(t, bc.v[j], mod[j]);
                if (x != 0) t = addm(t, vb.v[j], mod[j]);
                oa.v[j] = t;
                ob.v[j] = addm(vb.v[j], va.v[j], mod[j]);
            }
            if (x == 0 || p == 1) m.sm(iB, ob);  // x = L does not occur at p = 1
            if (x != 0 || p != 1) m.sm(iA, oa);
            if (x != 0 && p >= 2) m.sb(bidx, va);
            if (x == 0) {
                Gather<KC, L> g;
                g.init();
                gather_closures<KC, Idx, L>(T, W, p, hP, mainbase, sr, word, m, g, mod);
                g.finish(m, mod);
                if (p == 1) {
                    add_into<KC>(oa, g.acc, mod);
                    m.sm(iA, oa);
                } else {
                    m.sb(bidx, g.acc);
                }
            }
            return;
        }
        Gather<KC, L> g;
        g.init();
        if (x == 0) gather_closures<KC, Idx, L>(T, W, p, hP, mainbase, sr, word, m, g, mod);
        g.finish(m, mod);
        Vals<KC> oa, ob;
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            uint32_t t = va.v[j];
            if (a.read_blk) t = addm(t, bc.v[j], mod[j]);
            if (x != 0) t = addm(t, vb.v[j], mod[j]);
            else if (p == 1) t = addm(t, g.acc.v[j], mod[j]);
            oa.v[j] = t;
            ob.v[j] = addm(vb.v[j], va.v[j], mod[j]);
        }
        m.sm(iA, oa);
        if (x == 0 || p == 1) m.sm(iB, ob);  // x = L does not occur at p = 1
        if (p >= 2) {
#pragma unroll
            for (int j = 0; j < KC; ++j) ob.v[j] = (x == 0) ? g.acc.v[j] : va.v[j];
            m.sb(bidx, ob);
        }
        return;
    }
    if (x == 0) {
        // orbit NN -> LR
        const uint64_t iNN = (uint64_t)(mainbase + sr);
        const uint64_t iLR = (uint64_t)(mainbase + T.mcL(p, hP) + T.at(p - 1, hP + 1) + sr);
        Vals<KC> nn, lr, bc;
        if constexpr (V == 0) {
            m.lm(iNN, nn);
            m.lm(iLR, lr);
            if (a.read_blk) m.lb(bidx, bc);
        }
        Gather<KC, L> g;
        g.init();
        gather_closures<KC, Idx, L>(T, W, p, hP, mainbase, sr, word, m, g, mod);
        g.finish(m, mod);
        if constexpr (V == 1) {
            m.lm(iNN, nn);
            m.lm(iLR, lr);
            if (a.read_blk) m.lb(bidx, bc);
        }
        Vals<KC> onn, olr;
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            uint32_t t = a.read_blk ? addm(nn.v[j], bc.v[j], mod[j]) : nn.v[j];
            olr.v[j] = addm(lr.v[j], nn.v[j], mod[j]);
            if (p == 1) t = addm(t, g.acc.v[j], mod[j]);
            onn.v[j] = t;
        }
        m.sm(iNN, onn);
        m.sm(iLR, olr);
        if (p >= 2) m.sb(bidx, g.acc);
    } else if (x == 1) {
        // orbit RN -> NR (and at p=1 also NR -> RN)
        const uint64_t iNR = (uint64_t)(mainbase + T.at(p - 1, hP) + sr);
        const uint64_t iRN = (uint64_t)(mainbase + T.mat(p, hP) + sr);
        Vals<KC> nr, rn, bc, o;
        m.lm(iNR, nr);
        m.lm(iRN, rn);
        if (a.read_blk) m.lb(bidx, bc);
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            uint32_t t = addm(nr.v[j], rn.v[j], mod[j]);
            if (a.read_blk) t = addm(t, bc.v[j], mod[j]);
            o.v[j] = t;
        }
        m.sm(iNR, o);
        if (p >= 2) {
            m.sb(bidx, nr);
        } else {
#pragma unroll
            for (int j = 0; j < KC; ++j) o.v[j] = addm(rn.v[j], nr.v[j], mod[j]);
            m.sm(iRN, o);
        }
    } else {
        // orbit LN -> NL (cannot occur at p=1)
        const uint64_t iNL = (uint64_t)(mainbase + T.cL(p - 1, hP) + sr);
        const uint64_t iLN = (uint64_t)(mainbase + T.mcL(p, hP) + sr);
        Vals<KC> nl, ln, bc, o;
        m.lm(iNL, nl);
        m.lm(iLN, ln);
        if (a.read_blk) m.lb(bidx, bc);
#pragma unroll
        for (int j = 0; j < KC; ++j) {
            uint32_t t = addm(nl.v[j], ln.v[j], mod[j]);
            if (a.read_blk) t = addm(t, bc.v[j], mod[j]);
            o.v[j] = t;
        }
        m.sm(iNL, o);
        if (p >= 2) m.sb(bidx, nl);
    }
}

template <int KC, typename Idx, int L, int V, class Mem>
OFP_HD void cell_one(const CellLaunch& a, const Tab<Idx>& T, uint64_t bidx, const Mem& m) {
    cell_body<KC, Idx, L, V>(a, T, decode_blocked<Idx>(T, a.W, a.p, bidx), m);
}

// Packed capped table from a capped height table dc (kDpRows x kDpCols).
template <typename Idx>
inline void pack_pairs(const uint64_t (*dc)[kDpCols], Pair<Idx> (*q)[kDpCols]) {
    for (int pos = 0; pos < kDpRows; ++pos)
        for (int h = 0; h < kDpCols; ++h) {
            q[pos][h].n = (Idx)dc[pos][h];
            q[pos][h].c = (Idx)(dc[pos][h] + (h > 0 ? dc[pos][h - 1] : 0));
        }
}

// Packed table from the plain height table d (kDpRows x kDpCols).
template <typename Idx>
inline void pack_quads(const uint64_t (*d)[kDpCols], Quad<Idx> (*q)[kDpCols]) {
    for (int pos = 0; pos < kDpRows; ++pos)
        for (int h = 0; h < kDpCols; ++h) {
            Quad<Idx> e;
            const uint64_t up = pos + 1 < kDpRows ? d[pos + 1][h] : 0;
            const uint64_t up1 = (pos + 1 < kDpRows && h > 0) ? d[pos + 1][h - 1] : 0;
            e.zn = (Idx)d[pos][h];
            e.zc = (Idx)(d[pos][h] + (h > 0 ? d[pos][h - 1] : 0));
            e.mn = (Idx)up;
            e.mc = (Idx)(up + up1);
            q[pos][h] = e;
        }
}

// ---- occupancy groups ----
// A window of consecutive cells changes only its own positions and turns the
// R of a closure partner into L or back, which never changes whether a
// position is occupied (N or not).  So the words with one occupancy pattern of
// the positions outside the window (a group) are closed under the window's
// cells (reference: the bucket layout's LOW and HIGH windows,
// docs/b300-bucket-snake.md:7-20).  Two kinds, for a row split at `low`:
//   LOW groups, cells p = W-1..low+1: main and blocked words fix positions
//     0..low-1 (blocked positions below a cell are main positions);
//   HIGH groups, cells p = low..1: main words fix positions low+1..W-1, the
//     blocked words the same positions in their own numbering, low..W-2
//     (blocked position i above the cell is main position i+1).
// A group's words ranked in the global order (lexicographic, N < R < L from
// the top) have ranks that are sums of the same kind of entries as global
// ranks, taken from the group's completion table (one per numbering):
//   F[0][h] = [h == 0]
//   F[q][h] = F[q-1][h-1] + F[q-1][h+1]               position q-1 fixed occupied
//           = F[q-1][h]                               position q-1 fixed empty
//           = F[q-1][h] + F[q-1][h-1] + F[q-1][h+1]   position q-1 free
// with the entries of a fixed position counting only its allowed symbols:
//   occupied: zn = 0 (no N to pass), zc = F[pos][h-1] (the R completions)
//   empty:    zn = zc = F[pos][h] (every rank stays below zn: N is decoded)
// The packed table holds blocked-numbering entries in (zn, zc) and the main
// numbering's entries of the position one up in (mn, mc), which is what the
// cell body reads for main positions >= p (Tab::mat, Tab::mcL).  Below a
// cell both numberings agree.  Counts: main words Fm[W][1], blocked Fb[W-1][1].
struct Occupancy {
    int lo = 0, n = 0;  // fixed positions lo..lo+n-1
    uint32_t mask = 0;  // bit pos-lo set: position pos holds R or L; clear: N
    // -1: free, 0: fixed empty, 1: fixed occupied
    int at(int pos) const { return (pos < lo || pos >= lo + n) ? -1 : (int)((mask >> (pos - lo)) & 1u); }
};
// The two numberings of group `mask` of a row of width W split at `low`.
inline void group_occupancy(bool high, int W, int low, uint32_t mask, Occupancy& main_o, Occupancy& blk_o) {
    if (!high) {
        main_o = Occupancy{0, low, mask};
        blk_o = main_o;
    } else {
        main_o = Occupancy{low + 1, W - 1 - low, mask};
        blk_o = Occupancy{low, W - 1 - low, mask};
    }
}

inline void build_group_table(const Occupancy& o, uint64_t (*f)[kDpCols]) {
    for (int w = 0; w < kDpRows; ++w)
        for (int h = 0; h < kDpCols; ++h) f[w][h] = 0;
    f[0][0] = 1;
    for (int q = 1; q < kDpRows; ++q) {
        const int fix = o.at(q - 1);
        for (int h = 0; h < kDpCols; ++h) {
            const uint64_t dn = h > 0 ? f[q - 1][h - 1] : 0, up = h + 1 < kDpCols ? f[q - 1][h + 1] : 0;
            f[q][h] = fix == 1 ? dn + up : (fix == 0 ? f[q - 1][h] : f[q - 1][h] + dn + up);
        }
    }
}

// Rank contributions (zn, zc) of position pos at height h in a group table.
inline void group_entry(const Occupancy& o, const uint64_t (*f)[kDpCols], int pos, int h, uint64_t& zn,
                        uint64_t& zc) {
    if (pos >= kDpRows) {
        zn = zc = 0;
        return;
    }
    const uint64_t here = f[pos][h], below = h > 0 ? f[pos][h - 1] : 0;
    const int fix = o.at(pos);
    if (fix == 1) {
        zn = 0;
        zc = below;
    } else if (fix == 0) {
        zn = zc = here;
    } else {
        zn = here;
        zc = here + below;
    }
}

// Packed table of a group: (zn, zc) from the blocked numbering at pos, (mn, mc)
// from the main numbering at pos+1.
template <typename Idx>
inline void pack_group_quads(const Occupancy& ob, const uint64_t (*fb)[kDpCols], const Occupancy& om,
                             const uint64_t (*fm)[kDpCols], Quad<Idx> (*q)[kDpCols]) {
    for (int pos = 0; pos < kDpRows; ++pos)
        for (int h = 0; h < kDpCols; ++h) {
            uint64_t zn, zc, mn, mc;
            group_entry(ob, fb, pos, h, zn, zc);
            group_entry(om, fm, pos + 1, h, mn, mc);
            q[pos][h] = Quad<Idx>{(Idx)zn, (Idx)zc, (Idx)mn, (Idx)mc};
        }
}

}  // namespace cell
}  // namespace ofp
