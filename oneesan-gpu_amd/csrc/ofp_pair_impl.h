// ofp_pair_impl.h -- two cells per pass (the pair pass).
//
// A pair pass applies cells p and p-1 of a row to the whole state in place, so
// that every word crosses HBM once in and once out per two cells instead of once
// per cell.
//
// Unit of work: a *class* (P', d, S') -- P' = the content of main positions
// above p, S' = the content of positions below p-2 (rank sr among the suffixes
// of p-2 positions entered at height h+d), d = the height change over the
// window positions p, p-1, p-2 (-3..3), h = the height entering position p.
// The class holds the main words [P', a, b, c, S'] with dh(a)+dh(b)+dh(c) = d
// and the blocked words [P', u, v, S'] with dh(u)+dh(v) = d (blocked
// positions: P' above p-2, u at p-1, v at p-2).  Cells p and p-1 move values
// only inside a class, except for closures, whose sources flip one label
// outside the window and so lie in other classes.  The blocked words a class
// reads before cell p ([P', x, y, S'], x = position p-1) and writes after cell
// p-1 ([P', a, x', S'], a = position p, x' = position p-2) are the same index
// set, so each class owns its blocked words too.  One thread takes one class:
// it loads the class's words, applies both cells in registers and stores the
// words the two cells change (words they keep, such as [P', R, R, N, S'], are
// not stored; classes with d = +-3, [R, R, R] and [L, L, L] with their blocked
// [R, R] and [L, L], change nothing and get no launch).
//
// Closure sources of cell p are needed before cell p, those of cell p-1 after
// cell p; sources in other classes that cell p changes are rebuilt from their
// values before the pass by the cell-p rule that applies to them:
//   [P, N, R, R, S] (NR at cell p): NR + RN + blocked [P, R, R, S]
//   [P, N, L, L, S] (NL at cell p): NL + LN + blocked [P, L, L, S]
// (LR targets never occur among them: an RR source with an L at p would have
// ended the scan at that L), every other source is unchanged by cell p.
//
// Why in place works: only classes with d in {-1, 0, 1} have x=N words and so
// read other classes' words (closures); classes with d in {-3, -2, 2, 3} read
// only their own.  A reader's foreign reads are either words of classes with
// |d| >= 2 (the RR/LL sources and every word a rebuild needs: NR, RN, blocked
// [R, R] have d = -2, NL, LN, blocked [L, L] have d = +2) or words of other
// readers that the pass leaves unchanged ([P, R, R, L, S]: RR at cell p, RL at
// cell p-1; [P, R, L, L, S]: RL at cell p, LL at cell p-1).  So the readers
// run first (launches d = -1, 0, 1), while every |d| >= 2 class still holds
// its values from before the pass, and the |d| >= 2 classes after them.
//
// Classes are enumerated per window change d (one launch per d, so d is a
// compile-time constant and the words of a class are fixed register slots):
// class c of (p, d) is decoded by walking P' with completion counts
// G_d[m][h] = the (P', S') pairs below position m entered at height h,
// G_d[p][h] = D[p-2][h+d], G_d[m][h] = G_d[m-1][h] + G_d[m-1][h-1] + G_d[m-1][h+1].
//
// Reference semantics: the per-cell recurrence of
// src/common/gridfp_transition.hpp:113-146 (applied as in
// tests/factor_transfer_batch_test.cu:60-72); closure predecessors as
// src/common/gridfp_reverse.hpp:39-98.
#pragma once
#include "ofp_cell_impl.h"

namespace ofp {
namespace pair {

using cell::addm;
using cell::ldq;
using cell::ldz;
using cell::pick;
using cell::Quad;
using cell::Tab;
using cell::Vals;
using cell::Z2;

// symbols: 0 = N, 1 = R, 2 = L
OFP_HD constexpr int dh(int v) { return v == 2 ? 1 : (v == 1 ? -1 : 0); }

// Main window triples (a, b, c) (a at position p) with height change D, and
// blocked pairs (u, v) with height change D, in a fixed order.
template <int D>
struct Slots {
    static constexpr int count3() {
        int n = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int c = 0; c < 3; ++c) n += (dh(a) + dh(b) + dh(c) == D);
        return n;
    }
    static constexpr int count2() {
        int n = 0;
        for (int u = 0; u < 3; ++u)
            for (int v = 0; v < 3; ++v) n += (dh(u) + dh(v) == D);
        return n;
    }
    static constexpr int NM = count3();
    static constexpr int NB = count2();
    // slot of triple (a, b, c), -1 if its height change is not D
    static constexpr int m(int a, int b, int c) {
        int n = 0;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y)
                for (int z = 0; z < 3; ++z)
                    if (dh(x) + dh(y) + dh(z) == D) {
                        if (x == a && y == b && z == c) return n;
                        ++n;
                    }
        return -1;
    }
    static constexpr int b(int u, int v) {
        int n = 0;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y)
                if (dh(x) + dh(y) == D) {
                    if (x == u && y == v) return n;
                    ++n;
                }
        return -1;
    }
    static constexpr int tri(int i) {  // code a*9 + b*3 + c of slot i
        int n = 0;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y)
                for (int z = 0; z < 3; ++z)
                    if (dh(x) + dh(y) + dh(z) == D) {
                        if (n == i) return x * 9 + y * 3 + z;
                        ++n;
                    }
        return -1;
    }
    static constexpr int duo(int i) {  // code u*3 + v of slot i
        int n = 0;
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y)
                if (dh(x) + dh(y) == D) {
                    if (n == i) return x * 3 + y;
                    ++n;
                }
        return -1;
    }
};

// One pass: cells p and p-1 of an n-row, moduli off..off+kc-1.
struct PairLaunch {
    const uint32_t* main_in;
    const uint32_t* blk_in;
    uint32_t* main_out;
    uint32_t* blk_out;
    const void* dp;    // packed height table (cell::Quad, kDpRows x kDpCols)
    const void* ct;    // class table of (p, d): cell::Pair, kDpRows x kDpCols
    uint64_t classes;  // classes of (p, d)
    int W, p, d;
    int read_blk;      // blocked input of cell p is valid (false at the first cell of a row)
    int K, off, kc;    // interleave stride, first modulus, moduli in this launch
    int xcd_shift;
    int variant;        // kernel variant: bit 0 = gathers first (pair_body's GF), bit 1 = tighter registers
    uint32_t mod[kMaxKC];
};

// A decoded class.
template <typename Idx>
struct Class {
    Idx mainbase, blkbase;  // rank contributions of P' in the main / blocked numbering
    Idx sr;                 // rank of S' (suffix of p-2 positions entered at h + d)
    uint64_t word;          // P' symbols at bits 2*(m-1) for main positions m > p
    int h;                  // height entering main position p
};

// Decode class number c of (p, d): walk P' from the top position with the
// class completion counts ct (n = G_d[m-1][h], c = n + G_d[m-1][h-1]) and the
// height table T (the Quad at [m-1][h]: zn/zc = blocked contribution of main
// position m, mn/mc = its main contribution).
template <typename Idx>
OFP_HD Class<Idx> decode_class(const Tab<Idx>& T, const cell::Pair<Idx> (*ct)[cell::LdsCols<Idx>::value], int W,
                               int p, uint64_t c) {
    Class<Idx> k;
    Idx r = (Idx)c, mb = 0, bb = 0;
    uint64_t word = 0;
    int h = 1;
    for (int m = W - 1; m >= p + 1; --m) {
        const cell::Pair<Idx> g = ct[m][h];
        const Quad<Idx> e = ldq(T.q[m - 1][h]);
        const bool g0 = r >= g.n, g1 = r >= g.c;
        r -= g1 ? g.c : (g0 ? g.n : Idx(0));
        mb += g1 ? e.mc : (g0 ? e.mn : Idx(0));
        bb += g1 ? e.zc : (g0 ? e.zn : Idx(0));
        word = (word << 2) | (uint64_t)((int)g0 + (int)g1);
        h += g1 ? 1 : (g0 ? -1 : 0);
    }
    k.mainbase = mb;
    k.blkbase = bb;
    k.sr = r;
    k.word = word << (2 * p);
    k.h = h;
    return k;
}

// Rank contribution of symbol v at position pos entered at height h.
template <typename Idx>
OFP_HD Idx contrib(const Tab<Idx>& T, int pos, int v, int h) {
    if (v == 0) return Idx(0);
    const Z2<Idx> e = ldz(T.q[pos][h]);
    return v == 1 ? e.zn : e.zc;
}
// The same in the main numbering at main position pos >= 1 (the entries of the
// row below's main half, Tab::mat / Tab::mcL).  A pass reads main position p
// this way: with the global table and LOW groups both numberings give the same
// entry there, but a HIGH group of the split at p fixes blocked position p and
// leaves main position p free.  Below p the two numberings agree.
template <typename Idx>
OFP_HD Idx mcontrib(const Tab<Idx>& T, int pos, int v, int h) {
    if (v == 0) return Idx(0);
    const Z2<Idx> e = cell::ldm(T.q[pos - 1][h]);
    return v == 1 ? e.zn : e.zc;
}

// Memory of a pass (state-major, stride K, moduli off..off+KC-1).  Loads and
// stores may address the same arrays (the pass runs in place), so no pointer
// here is declared non-aliasing.
template <int KC, typename Idx>
struct Io {
    const uint32_t* mi;
    const uint32_t* bi;
    uint32_t* mo;
    uint32_t* bo;
    int K, off;
    OFP_HD void lm(uint64_t i, Vals<KC>& v) const {
        cell::load_vals<KC, cell::kAoS>(mi, (uint64_t)(Idx)i, K, off, v);
    }
    OFP_HD void lb(uint64_t i, Vals<KC>& v) const {
        cell::load_vals<KC, cell::kAoS>(bi, (uint64_t)(Idx)i, K, off, v);
    }
    OFP_HD void sm(uint64_t i, const Vals<KC>& v) const {
        cell::store_vals<KC, cell::kAoS>(mo, (uint64_t)(Idx)i, K, off, v);
    }
    OFP_HD void sb(uint64_t i, const Vals<KC>& v) const {
        cell::store_vals<KC, cell::kAoS>(bo, (uint64_t)(Idx)i, K, off, v);
    }
};

template <int KC>
OFP_HD void vzero(Vals<KC>& v) {
#pragma unroll
    for (int j = 0; j < KC; ++j) v.v[j] = 0;
}
template <int KC>
OFP_HD void vadd(Vals<KC>& acc, const Vals<KC>& x, const uint32_t* mod) {
#pragma unroll
    for (int j = 0; j < KC; ++j) acc.v[j] = addm(acc.v[j], x.v[j], mod[j]);
}

// Closure sources of the x=N word of cell p-1 in other classes, at their value
// after cell p, added into acc (their classes have |d| >= 2 or leave them
// unchanged, see the head of this file).  u = the symbol at main position p (the lowest
// of the cell-(p-1) prefix P'' = (P', u)), hq = h + dh(u).
//   LL: [P', u, L, L, S*] (S* = S' with a top-level L turned into R), scan of S';
//   RR: [P'*, u, R, R, S'] (P'* = P' with a top-level R turned into L), scan of
//       P' (the in-class source with u itself flipped is added by the caller).
// u = N: the source was an NL / NR target of cell p and is rebuilt from old
// (its blocked term only when the blocked input of cell p is valid).
template <int KC, typename Idx>
OFP_HD void closures_after(const Tab<Idx>& T, int W, int p, const Class<Idx>& k, int u, bool read_blk,
                           const Io<KC, Idx>& io, Vals<KC>& acc, const uint32_t* mod) {
    const int h = k.h, hq = h + dh(u);
    const Idx mbq = k.mainbase + mcontrib(T, p, u, h);  // main contribution of P''
    Vals<KC> t;
    // LL (cell p-1 >= 3): scan S' downwards from its rank; top level = height hq
    if (p - 1 >= 3) {
        const Idx baseLL = mbq + T.cL(p - 1, hq) + T.cL(p - 2, hq + 1);
        // u = N: LN = NL - cL(p-1, h) + cL(p, h); blocked [P', L, L, S*] = NL - mainbase + blkbase
        const Idx dLN = T.mcL(p, h) - T.cL(p - 1, h);
        const Idx dBL = k.blkbase - k.mainbase;
        Idx rr = k.sr, s2 = 0;
        int hh = hq;
        for (int pos = p - 3; pos >= 0; --pos) {
            const Z2<Idx> e = ldz(T.q[pos][hh]);
            const Z2<Idx> e2 = ldz(T.q[pos][hh + 2]);
            const bool g0 = rr >= e.zn, g1 = rr >= e.zc;
            rr -= g1 ? e.zc : (g0 ? e.zn : Idx(0));
            if (g1 && hh == hq) {
                const Idx i = baseLL + s2 + e2.zn + rr;
                io.lm((uint64_t)i, t);
                vadd<KC>(acc, t, mod);
                if (u == 0) {
                    io.lm((uint64_t)(Idx)(i + dLN), t);
                    vadd<KC>(acc, t, mod);
                    if (read_blk) {
                        io.lb((uint64_t)(Idx)(i + dBL), t);
                        vadd<KC>(acc, t, mod);
                    }
                }
            }
            if (g0 && !g1 && hh == hq) break;
            s2 += g1 ? e2.zc : (g0 ? e2.zn : Idx(0));
            hh += g1 ? 1 : (g0 ? -1 : 0);
        }
    }
    // RR over P' (positions p+1 and up); an L at the top level ends it.  u = L
    // ends it at once (its own position is the top level), u = R raises the
    // depth of everything above it.
    if (u == 2) return;
    {
        // candidate main index = ub + (L contribution at the flip) + u2:
        //   ub: main contributions of the unscanned part of P', plus u at its
        //       height in P'* (h + 2), the two R's of the window and S';
        //   u2: main contributions of the scanned part of P' two levels higher.
        // The blocked index of [P'*, R, R, S'] (needed when u = N) differs from
        // the main index of [P'*, N, R, R, S'] by the blocked-minus-main
        // contributions of P'*: db (unscanned) + dl (flip) + d2 (scanned, +2).
        const int hs = h + 2;  // height entering position p in P'*
        Idx ub = k.mainbase + mcontrib(T, p, u, hs) + T.at(p - 1, hq + 2) + T.at(p - 2, hq + 1) + k.sr;
        Idx u2 = 0, db = k.blkbase - k.mainbase, d2 = 0;
        const Idx dRN = T.mat(p, hs) - T.at(p - 1, hs);  // u = N: RN - NR in P'*
        int Ha = h;  // height below the scanned position (after u and the window)
        uint64_t w = k.word >> (2 * p);
        for (int mp = p + 1; mp <= W - 1; ++mp) {
            const uint32_t v = (uint32_t)w & 3u;
            w >>= 2;
            const bool isR = v == 1, isL = v == 2;
            const int H = Ha + (isR ? 1 : (isL ? -1 : 0));
            const Quad<Idx> e0 = ldq(T.q[mp - 1][H]);
            const Quad<Idx> e2 = ldq(T.q[mp - 1][H + 2]);
            // remove this position's contributions from the unscanned sums
            ub -= isL ? e0.mc : (isR ? e0.mn : Idx(0));
            db -= isL ? (e0.zc - e0.mc) : (isR ? (e0.zn - e0.mn) : Idx(0));
            const bool atTop = Ha == hq;  // top level of the x'=N word's prefix
            if (isR && atTop) {
                const Idx i = ub + e0.mc + u2;
                io.lm((uint64_t)i, t);
                vadd<KC>(acc, t, mod);
                if (u == 0) {
                    io.lm((uint64_t)(Idx)(i + dRN), t);
                    vadd<KC>(acc, t, mod);
                    if (read_blk) {
                        io.lb((uint64_t)(Idx)(i + db + (e0.zc - e0.mc) + d2), t);
                        vadd<KC>(acc, t, mod);
                    }
                }
            }
            if (isL && atTop) break;
            u2 += isL ? e2.mc : (isR ? e2.mn : Idx(0));
            d2 += isL ? (e2.zc - e2.mc) : (isR ? (e2.zn - e2.mn) : Idx(0));
            Ha = H;
        }
    }
}

// The pass over one class of window change D.  GF = 0: the class's words are
// loaded first and the closure scans run while they are in flight; GF = 1: the
// closure sums (the parts read from memory) are gathered first and the class's
// words loaded after them, so that fewer values are live during the scans.
template <int KC, typename Idx, int D, int GF = 0>
OFP_HD void pair_body(const PairLaunch& a, const Tab<Idx>& T, const Class<Idx>& k, const Io<KC, Idx>& io) {
    using S = Slots<D>;
    const int W = a.W, p = a.p, h = k.h;
    uint32_t mod[KC];
#pragma unroll
    for (int j = 0; j < KC; ++j) mod[j] = a.mod[j];

    constexpr int yN = D == 0 ? 0 : (D == -1 ? 1 : (D == 1 ? 2 : -1));  // dh(yN) = D
    // indices and presence of the class's words
    uint64_t im[S::NM > 0 ? S::NM : 1];
    bool pm[S::NM > 0 ? S::NM : 1];
#pragma unroll
    for (int i = 0; i < S::NM; ++i) {
        const int code = S::tri(i), x = code / 9, y = (code / 3) % 3, z = code % 3;
        const int h1 = h + dh(x), h2 = h1 + dh(y);
        pm[i] = h1 >= 0 && h2 >= 0;
        im[i] = (uint64_t)(Idx)(k.mainbase + mcontrib(T, p, x, h) + contrib(T, p - 1, y, h1 < 0 ? 0 : h1) +
                                contrib(T, p - 2, z, h2 < 0 ? 0 : h2) + k.sr);
    }
    uint64_t ib[S::NB > 0 ? S::NB : 1];
    bool pb[S::NB > 0 ? S::NB : 1];
#pragma unroll
    for (int i = 0; i < S::NB; ++i) {
        const int code = S::duo(i), x = code / 3, y = code % 3;
        const int h1 = h + dh(x);
        pb[i] = h1 >= 0;
        ib[i] = (uint64_t)(Idx)(k.blkbase + contrib(T, p - 1, x, h) + contrib(T, p - 2, y, h1 < 0 ? 0 : h1) + k.sr);
    }
    // closure sums read from memory: cell p's x=N word (sources before cell p)
    // and cell p-1's x'=N word (sources in other classes, rebuilt as after cell p)
    Vals<KC> clo, clo2;
    vzero<KC>(clo);
    vzero<KC>(clo2);
    auto gather_p = [&]() {
        if constexpr (yN >= 0) {
            if (pb[S::b(0, yN)]) {
                cell::Gather<KC, cell::kAoS> g;
                g.init();
                const cell::Direct<KC, cell::kAoS, Idx> old{const_cast<uint32_t*>(io.mi),
                                                            const_cast<uint32_t*>(io.bi), io.K, io.off};
                const Idx srp = contrib(T, p - 2, yN, h) + k.sr;  // rank of S = (y, S') entered at h
                cell::gather_closures<KC, Idx, cell::kAoS>(T, W, p, h, k.mainbase, srp, k.word, old, g, mod);
                g.finish(old, mod);
                clo = g.acc;
            }
        }
    };
    auto gather_p1 = [&]() {
        if constexpr (yN >= 0) {
            if (pb[S::b(yN, 0)]) closures_after<KC, Idx>(T, W, p, k, yN, a.read_blk != 0, io, clo2, mod);
        }
    };
    if constexpr (GF == 1) {
        gather_p();
        gather_p1();
    }
    Vals<KC> m[S::NM > 0 ? S::NM : 1];
#pragma unroll
    for (int i = 0; i < S::NM; ++i) {
        if (pm[i]) io.lm(im[i], m[i]);
        else vzero<KC>(m[i]);
    }
    Vals<KC> bl[S::NB > 0 ? S::NB : 1];
#pragma unroll
    for (int i = 0; i < S::NB; ++i) {
        if (pb[i] && a.read_blk) io.lb(ib[i], bl[i]);
        else vzero<KC>(bl[i]);
    }

    // Cell-by-cell, the orbits whose blocked word is x=R or x=L first and the x=N
    // orbit (whose blocked output is the closure sum) last, so that the closure
    // scans run with as few values live as possible; bl[i] holds, in turn, the
    // blocked word before cell p, after cell p and after cell p-1.  A word is
    // stored as soon as it is final.
    //
    // A word [P', a, b, c, S'] keeps its value when neither (a, b) is a target of
    // cell p nor (b, c) one of cell p-1 (targets: NN, LR, NR, NL; at cell 1 also
    // RN): RRN, RRR, RRL, LLN, LLL, RLN, RLL.  Those are not stored.

    // ---- cell p: orbits of the blocked words (x, y) (x at p-1, y at p-2) ----
#pragma unroll
    for (int i = 0; i < S::NB; ++i) {
        const int code = S::duo(i), x = code / 3, y = code % 3;
        if (x == 1) {
            const int iNR = S::m(0, 1, y), iRN = S::m(1, 0, y);
            const Vals<KC> nr = m[iNR];
            vadd<KC>(m[iNR], m[iRN], mod);
            vadd<KC>(m[iNR], bl[i], mod);
            bl[i] = nr;
        } else if (x == 2) {
            const int iNL = S::m(0, 2, y), iLN = S::m(2, 0, y);
            const Vals<KC> nl = m[iNL];
            vadd<KC>(m[iNL], m[iLN], mod);
            vadd<KC>(m[iNL], bl[i], mod);
            bl[i] = nl;
        }
    }
    if constexpr (yN >= 0) {
        constexpr int i = S::b(0, yN);
        constexpr int iNN = S::m(0, 0, yN), iLR = S::m(2, 1, yN);
        vadd<KC>(m[iLR], m[iNN], mod);
        vadd<KC>(m[iNN], bl[i], mod);
        if constexpr (GF == 0) gather_p();
        bl[i] = clo;  // blocked output: the closure sum
    }

    // ---- cell p-1: orbits of the blocked words (u, x') (u at p, x' at p-2) ----
    auto store_main = [&](int i) {
        if (pm[i]) io.sm(im[i], m[i]);
    };
    auto store_blk = [&](int i) {
        if (p - 1 >= 2 && pb[i]) io.sb(ib[i], bl[i]);
    };
#pragma unroll
    for (int i = 0; i < S::NB; ++i) {
        const int code = S::duo(i), u = code / 3, x = code % 3;
        if (x == 1) {
            const int iNR = S::m(u, 0, 1), iRN = S::m(u, 1, 0);
            const Vals<KC> nr = m[iNR];
            vadd<KC>(m[iNR], m[iRN], mod);
            vadd<KC>(m[iNR], bl[i], mod);
            bl[i] = nr;
            if (p - 1 == 1) vadd<KC>(m[iRN], nr, mod);
            store_main(iNR);
            // RN is a target of cell p when u = N... it is [P', u, R, N]: (u, R) at
            // cell p is NR (u = N) or LR (u = L), changed; RRN keeps (cell 1 aside)
            if (u != 1 || p - 1 == 1) store_main(iRN);
            store_blk(i);
        } else if (x == 2) {
            const int iNL = S::m(u, 0, 2), iLN = S::m(u, 2, 0);
            const Vals<KC> nl = m[iNL];
            vadd<KC>(m[iNL], m[iLN], mod);
            vadd<KC>(m[iNL], bl[i], mod);
            bl[i] = nl;
            store_main(iNL);
            // LN = [P', u, L, N]: (u, L) at cell p is NL (u = N, changed), RL or LL (keep)
            if (u == 0) store_main(iLN);
            store_blk(i);
        }
    }
    if constexpr (yN >= 0) {
        constexpr int uN = yN;  // the x'=N word has dh(u) = D
        constexpr int i = S::b(uN, 0);
        constexpr int iNN = S::m(uN, 0, 0), iLR = S::m(uN, 2, 1);
        if (pb[i]) {
            // in-class sources, after cell p: RL = [P', u, R, L, S'] (cell p-1 >= 2),
            // and with u = R the source with u itself turned into L: [P', L, R, R, S']
            // (neither is a target of cell p-1)
            if (p - 1 >= 2) {
                constexpr int iRL = S::m(uN, 1, 2);
                if (iRL >= 0 && pm[iRL]) vadd<KC>(clo2, m[iRL], mod);
            }
            if constexpr (uN == 1) {
                constexpr int iLRR = S::m(2, 1, 1);
                if (pm[iLRR]) vadd<KC>(clo2, m[iLRR], mod);
            }
            if constexpr (GF == 0) gather_p1();
        }
        vadd<KC>(m[iLR], m[iNN], mod);
        vadd<KC>(m[iNN], bl[i], mod);
        if (p - 1 == 1) vadd<KC>(m[iNN], clo2, mod);
        bl[i] = clo2;
        store_main(iNN);
        store_main(iLR);
        store_blk(i);
    }
    // words outside every cell-(p-1) orbit ((b, c) = RR, LL or RL) that cell p
    // changed: (a, b) = NR, NL or LR, i.e. NRR, NRL, NLL, LRR, LRL
#pragma unroll
    for (int i = 0; i < S::NM; ++i) {
        const int code = S::tri(i);
        if (code == 0 * 9 + 1 * 3 + 1 || code == 0 * 9 + 1 * 3 + 2 || code == 0 * 9 + 2 * 3 + 2 ||
            code == 2 * 9 + 1 * 3 + 1 || code == 2 * 9 + 1 * 3 + 2)
            store_main(i);
    }
}

// Class completion table of (p, d) from the plain height table d0
// (kDpRows x kDpCols): out[m][h] = {G[m-1][h], G[m-1][h] + G[m-1][h-1]} for
// m in p+1..kDpRows-1 (other rows zero).  Returns G[W-1][1] for W = kDpRows-1
// unused; the class count of an n-row is g[W-1][1], see class_count().
inline void build_class_table(const uint64_t (*d0)[kDpCols], int p, int d, uint64_t (*g)[kDpCols]) {
    for (int m = 0; m < kDpRows; ++m)
        for (int h = 0; h < kDpCols; ++h) g[m][h] = 0;
    for (int h = 0; h < kDpCols; ++h) {
        const int e = h + d;
        g[p][h] = (p >= 2 && e >= 0 && e < kDpCols) ? d0[p - 2][e] : 0;
    }
    for (int m = p + 1; m < kDpRows; ++m)
        for (int h = 0; h < kDpCols; ++h)
            g[m][h] = g[m - 1][h] + (h > 0 ? g[m - 1][h - 1] : 0) + (h + 1 < kDpCols ? g[m - 1][h + 1] : 0);
}

}  // namespace pair
}  // namespace ofp

namespace ofp {
namespace pair {

// Packed class table of (p, d) with C columns per row: row m holds
// {G_d[m-1][h], G_d[m-1][h] + G_d[m-1][h-1]} for m > p (rows <= p are zero);
// C = cell::LdsCols<Idx>::value, the layout the kernels keep in LDS.
template <typename Idx, int C>
inline void pack_class_table(const uint64_t (*d0)[kDpCols], int p, int d, cell::Pair<Idx> (*out)[C]) {
    uint64_t g[kDpRows][kDpCols];
    build_class_table(d0, p, d, g);
    for (int m = 0; m < kDpRows; ++m)
        for (int h = 0; h < C; ++h) {
            out[m][h].n = 0;
            out[m][h].c = 0;
            if (m > p && m >= 1) {
                out[m][h].n = (Idx)g[m - 1][h];
                out[m][h].c = (Idx)(g[m - 1][h] + (h > 0 ? g[m - 1][h - 1] : 0));
            }
        }
}

// Classes of (p, d) in an n-row: G_d[W-1][1].
inline uint64_t class_count(const uint64_t (*d0)[kDpCols], int W, int p, int d) {
    uint64_t g[kDpRows][kDpCols];
    build_class_table(d0, p, d, g);
    return W - 1 >= p ? g[W - 1][1] : 0;
}

// ---- pair passes over one occupancy group (cell::Occupancy, ofp_cell_impl.h) ----
// The classes of a group are those of the whole state whose words lie in the
// group: S' takes the group's completion counts below the window, G[p][h] =
// F[p-2][h+d] (F: the group's main-numbering completion table; below a cell
// both numberings agree), and each position m of P' above p obeys the group's
// constraint there (fixed occupied: R or L, fixed empty: N, free: any):
//   G[m][h] = G[m-1][h-1] + G[m-1][h+1]                (occupied)
//           = G[m-1][h]                                (empty)
//           = G[m-1][h] + G[m-1][h-1] + G[m-1][h+1]    (free)
// and the packed decode entry of position m counts only the symbols allowed
// there (as cell::group_entry).  The window positions p, p-1, p-2 of a pair
// pass must be free (a LOW group fixes positions below low, a HIGH group above
// low; the passes stay inside their half), so the class's words are the same
// slots as in the whole state.  With the global table and no fixed position
// this is build_class_table.
inline void build_group_class_table(const cell::Occupancy& om, const uint64_t (*f)[kDpCols], int p, int d,
                                    uint64_t (*g)[kDpCols]) {
    for (int m = 0; m < kDpRows; ++m)
        for (int h = 0; h < kDpCols; ++h) g[m][h] = 0;
    for (int h = 0; h < kDpCols; ++h) {
        const int e = h + d;
        g[p][h] = (p >= 2 && e >= 0 && e < kDpCols) ? f[p - 2][e] : 0;
    }
    for (int m = p + 1; m < kDpRows; ++m) {
        const int fix = om.at(m);
        for (int h = 0; h < kDpCols; ++h) {
            const uint64_t dn = h > 0 ? g[m - 1][h - 1] : 0, up = h + 1 < kDpCols ? g[m - 1][h + 1] : 0;
            g[m][h] = fix == 1 ? dn + up : (fix == 0 ? g[m - 1][h] : g[m - 1][h] + dn + up);
        }
    }
}

template <typename Idx, int C>
inline void pack_group_class_table(const cell::Occupancy& om, const uint64_t (*f)[kDpCols], int p, int d,
                                   cell::Pair<Idx> (*out)[C]) {
    uint64_t g[kDpRows][kDpCols];
    build_group_class_table(om, f, p, d, g);
    for (int m = 0; m < kDpRows; ++m)
        for (int h = 0; h < C; ++h) {
            out[m][h].n = 0;
            out[m][h].c = 0;
            if (m > p && m >= 1) {
                const uint64_t here = g[m - 1][h], below = h > 0 ? g[m - 1][h - 1] : 0;
                const int fix = om.at(m);
                out[m][h].n = (Idx)(fix == 1 ? 0 : here);
                out[m][h].c = (Idx)(fix == 1 ? below : (fix == 0 ? here : here + below));
            }
        }
}

inline uint64_t group_class_count(const cell::Occupancy& om, const uint64_t (*f)[kDpCols], int W, int p, int d) {
    uint64_t g[kDpRows][kDpCols];
    build_group_class_table(om, f, p, d, g);
    return W - 1 >= p ? g[W - 1][1] : 0;
}

}  // namespace pair
}  // namespace ofp
