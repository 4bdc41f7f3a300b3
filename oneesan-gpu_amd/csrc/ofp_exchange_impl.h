// ofp_exchange_impl.h -- the exchange between the two placements of a row.
//
// A row of width W split at `low` runs its high cells W-1..low+1 over LOW
// groups (words with one occupancy of positions 0..low-1) and its low cells
// low..1 over HIGH groups (one occupancy of main positions low+1..W-1; blocked
// words: blocked positions low..W-2).  With the state owner-partitioned by
// group over several GPUs, every word of a LOW window moves, between the two
// halves, to the GPU that owns its HIGH group: the per-row all-to-all of the
// reference's 8x8 bucket transpose (docs/b300-bucket-snake.md:7-20, the
// H-major / L-major placements).  This header holds the per-word index math,
// for host and device:
//
//   * decode: the symbols of the word with local rank r in its LOW group, by a
//     walk over the group's packed table (the same table the cell kernel uses);
//   * destination: the word's HIGH group mask (occupancy of the fixed
//     positions base..base+nf-1, base = low+1 for main words, low for blocked
//     words) and its rank inside that group, a sum of rank contributions:
//       free position pos < base: the plain height table's zn (R) / zc (L);
//       fixed position pos = base+j: R and N contribute 0, an L contributes
//       F[pos][h-1], where F is the group's completion table
//         F[base][h] = D[base][h],
//         F[q+1][h]  = F[q][h-1] + F[q][h+1]   (position q occupied)
//                    = F[q][h]                 (position q empty).
//     F[base+j] depends only on the mask bits below j, so one table holds
//     every group's rows: row (2^j - 1) + (mask & (2^j - 1)), 2^nf - 1 rows
//     (1 MB at nf = 13, kDpCols 32-bit columns).
// Ranks are summed mod 2^32: every HIGH group of the W = 28 split holds fewer
// than 2^32 words, so the true rank comes out whatever the unused entries hold.
#pragma once
#include <functional>
#include <vector>

#include "ofp_cell_impl.h"

namespace ofp {
namespace xchg {

using cell::Quad;
using cell::Z2;

// Symbols (2 bits per position, bit 2*pos) of the word with local rank r in a
// group with packed table q; positions top..stop are decoded (main words: top
// = W-1 from the main entries, blocked words: top = W-2 from the blocked
// entries), the rest stay 0.
template <bool Main>
OFP_HD uint64_t decode_group_word(const Quad<uint32_t> (*q)[kDpCols], int top, int stop, uint32_t r) {
    uint64_t word = 0;
    int h = 1;
    for (int pos = top; pos >= stop; --pos) {
        uint32_t zn, zc;
        if (Main && pos >= 1) {
            const Z2<uint32_t> e = cell::ldm(q[pos - 1][h]);
            zn = e.zn;
            zc = e.zc;
        } else {
            const Z2<uint32_t> e = cell::ldz(q[pos][h]);
            zn = e.zn;
            zc = e.zc;
        }
        const bool g0 = r >= zn, g1 = r >= zc;
        r -= g1 ? zc : (g0 ? zn : 0u);
        word |= (uint64_t)((int)g0 + (int)g1) << (2 * pos);
        h += g1 ? 1 : (g0 ? -1 : 0);
    }
    return word;
}

// Occupancy mask of positions base..base+nf-1 of a word.
OFP_HD uint32_t high_mask(uint64_t word, int base, int nf) {
    uint32_t m = 0;
    for (int j = 0; j < nf; ++j) m |= (uint32_t)(((word >> (2 * (base + j))) & 3u) != 0) << j;
    return m;
}

// Rank of a word (positions top..0, entered at height 1) inside its HIGH group
// `mask`: ht = the group-row table described at the head of this file, d = the
// plain packed height table (zn, zc = D entries).
OFP_HD uint32_t high_rank(uint64_t word, uint32_t mask, int top, int base, const uint32_t* ht,
                          const Quad<uint32_t> (*d)[kDpCols]) {
    uint32_t rank = 0;
    int h = 1;
    for (int pos = top; pos >= 0; --pos) {
        const int v = (int)((word >> (2 * pos)) & 3u);
        if (pos >= base) {
            const int j = pos - base;
            if (v == 2 && h >= 1) rank += ht[(size_t)((1u << j) - 1u + (mask & ((1u << j) - 1u))) * kDpCols + (h - 1)];
        } else {
            const Z2<uint32_t> e = cell::ldz(d[pos][h]);
            rank += v == 2 ? e.zc : (v == 1 ? e.zn : 0u);
        }
        h += v == 2 ? 1 : (v == 1 ? -1 : 0);
    }
    return rank;
}

// decode_group_word and high_mask / high_rank in one walk from the top: the
// fixed HIGH positions first (their L's and the heights there are recorded,
// since the group row of position base+j depends on the mask bits below j, not
// yet decoded), then the free positions, whose plain-table contributions are
// summed as they are decoded; the recorded L's are added last.  A legal word
// of W <= 28 positions stays at height <= 14, so h-1 fits 4 bits per position.
template <bool Main>
OFP_HD uint32_t decode_rank(const Quad<uint32_t> (*q)[kDpCols], const Quad<uint32_t> (*d)[kDpCols],
                            const uint32_t* ht, int top, int base, uint32_t r, uint64_t& word_out,
                            uint32_t& mask_out) {
    uint64_t word = 0, hpack = 0;
    uint32_t mask = 0, lmask = 0, rank = 0;
    int h = 1;
    for (int pos = top; pos >= base; --pos) {
        const Z2<uint32_t> e = (Main && pos >= 1) ? cell::ldm(q[pos - 1][h]) : cell::ldz(q[pos][h]);
        const bool g0 = r >= e.zn, g1 = r >= e.zc;
        r -= g1 ? e.zc : (g0 ? e.zn : 0u);
        const int j = pos - base;
        mask |= (uint32_t)g0 << j;
        if (g1 && h >= 1) {
            lmask |= 1u << j;
            hpack |= (uint64_t)(h - 1) << (4 * j);
        }
        word |= (uint64_t)((int)g0 + (int)g1) << (2 * pos);
        h += g1 ? 1 : (g0 ? -1 : 0);
    }
    for (int pos = base - 1; pos >= 0; --pos) {
        const Z2<uint32_t> e = (Main && pos >= 1) ? cell::ldm(q[pos - 1][h]) : cell::ldz(q[pos][h]);
        const Z2<uint32_t> f = cell::ldz(d[pos][h]);
        const bool g0 = r >= e.zn, g1 = r >= e.zc;
        r -= g1 ? e.zc : (g0 ? e.zn : 0u);
        rank += g1 ? f.zc : (g0 ? f.zn : 0u);
        word |= (uint64_t)((int)g0 + (int)g1) << (2 * pos);
        h += g1 ? 1 : (g0 ? -1 : 0);
    }
    while (lmask) {
        const int j = __builtin_ctz(lmask);
        lmask &= lmask - 1u;
        const uint32_t below = mask & ((1u << j) - 1u);
        rank += ht[(size_t)((1u << j) - 1u + below) * kDpCols + (uint32_t)((hpack >> (4 * j)) & 15u)];
    }
    word_out = word;
    mask_out = mask;
    return rank;
}

// The top part of a word alone: positions top..base decoded from local rank r
// (the walk of decode_rank), giving the rank r_rest of the bottom part among the
// group's bottoms entered at height h, the HIGH mask, and the fixed positions'
// recorded L's (lmask, hpack).  Every word with the same top part lies in one
// run of consecutive local ranks (a block: r - r_rest .. + bottoms(h) - 1) and
// shares mask, share and the top part's HIGH-rank contribution (top_high).
struct Top {
    uint32_t r_rest, mask, lmask;
    uint64_t hpack;
    int h;
};
template <bool Main>
OFP_HD Top decode_top(const Quad<uint32_t> (*q)[kDpCols], int top, int base, uint32_t r) {
    Top t{};
    int h = 1;
    for (int pos = top; pos >= base; --pos) {
        const Z2<uint32_t> e = (Main && pos >= 1) ? cell::ldm(q[pos - 1][h]) : cell::ldz(q[pos][h]);
        const bool g0 = r >= e.zn, g1 = r >= e.zc;
        r -= g1 ? e.zc : (g0 ? e.zn : 0u);
        const int j = pos - base;
        t.mask |= (uint32_t)g0 << j;
        if (g1 && h >= 1) {
            t.lmask |= 1u << j;
            t.hpack |= (uint64_t)(h - 1) << (4 * j);
        }
        h += g1 ? 1 : (g0 ? -1 : 0);
    }
    t.r_rest = r;
    t.h = h;
    return t;
}
// The fixed positions' part of the HIGH rank (the recorded L's).
OFP_HD uint32_t top_high(const Top& t, const uint32_t* ht) {
    uint32_t rank = 0, lm = t.lmask;
    while (lm) {
        const int j = __builtin_ctz(lm);
        lm &= lm - 1u;
        rank += ht[(size_t)((1u << j) - 1u + (t.mask & ((1u << j) - 1u))) * kDpCols +
                   (uint32_t)((t.hpack >> (4 * j)) & 15u)];
    }
    return rank;
}

// The bottoms of a LOW group (positions base-1..0, entered at height h, the
// group's occupancy on its fixed positions) in the group's order, and for each
// its rank contribution in the plain numbering (what a HIGH group, free below
// its fixed block, counts for it): lut[off[h] + j] for the j-th bottom entered
// at h; off has kDpCols + 1 entries, off[h+1] - off[h] = f[base][h] bottoms.
inline void build_bottom_lut(const cell::Occupancy& o, const uint64_t (*f)[kDpCols], const uint64_t (*d)[kDpCols],
                             int base, std::vector<uint32_t>& lut, uint32_t* off) {
    lut.clear();
    for (int h0 = 0; h0 < kDpCols; ++h0) {
        off[h0] = (uint32_t)lut.size();
        if (base == 0) {
            if (h0 == 0) lut.push_back(0);
            continue;
        }
        if (f[base][h0] == 0) continue;
        // depth-first in order N < R < L at each position from base-1 down
        std::function<void(int, int, uint32_t)> rec = [&](int pos, int h, uint32_t rank) {
            if (pos < 0) {
                if (h == 0) lut.push_back(rank);
                return;
            }
            const int fix = o.at(pos);
            for (int v = 0; v < 3; ++v) {
                if (fix >= 0 && (v != 0) != (fix == 1)) continue;
                const int nh = h + (v == 2) - (v == 1);
                if (nh < 0 || nh >= kDpCols || f[pos][nh] == 0) continue;
                const uint64_t here = d[pos][h], below = h > 0 ? d[pos][h - 1] : 0;
                const uint32_t c = v == 0 ? 0u : (uint32_t)(v == 1 ? here : here + below);
                rec(pos - 1, nh, rank + c);
            }
        };
        rec(base - 1, h0, 0u);
    }
    off[kDpCols] = (uint32_t)lut.size();
}

// Share (destination shard) of a HIGH mask: shares own consecutive runs of
// masks, share s the masks first[s]..first[s+1]-1.
OFP_HD int share_of(uint32_t mask, const uint32_t* first, int shares) {
    int s = 0;
    for (int i = 1; i < shares; ++i) s += mask >= first[i];
    return s;
}

// The group-row table for fixed positions base..base+nf-1 (rows as described
// at the head of this file), from the plain height table d.
inline void build_high_rows(const uint64_t (*d)[kDpCols], int base, int nf, uint32_t* ht) {
    std::vector<uint64_t> rows((size_t)((1u << nf) - 1u) * kDpCols, 0);
    for (int h = 0; h < kDpCols; ++h) rows[h] = d[base][h];
    for (int j = 1; j < nf; ++j) {
        const uint32_t below = (1u << (j - 1)) - 1u, here = (1u << j) - 1u;
        for (uint32_t low = 0; low < (1u << j); ++low) {
            const uint64_t* prev = &rows[(size_t)(below + (low & ((1u << (j - 1)) - 1u))) * kDpCols];
            uint64_t* cur = &rows[(size_t)(here + low) * kDpCols];
            const bool occupied = (low >> (j - 1)) & 1u;
            for (int h = 0; h < kDpCols; ++h)
                cur[h] = occupied ? (h > 0 ? prev[h - 1] : 0) + (h + 1 < kDpCols ? prev[h + 1] : 0) : prev[h];
        }
    }
    for (size_t i = 0; i < rows.size(); ++i) ht[i] = (uint32_t)rows[i];
}

}  // namespace xchg
}  // namespace ofp
