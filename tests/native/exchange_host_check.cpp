// Host check of the exchange's index math (oneesan-gpu_amd/csrc/
// ofp_exchange_impl.h): decode_group_word, high_mask, high_rank, the fused
// decode_rank, build_high_rows, and the block route (decode_top, top_high and
// the group's bottom table from build_bottom_lut), built for the CPU.
//
// `exchange_host_check small NMAX`: for every n <= NMAX and every split low,
// every main and blocked word is enumerated in global rank order (the oracle's
// or_unrank); its LOW group and rank inside the LOW group, and its HIGH group
// and rank inside the HIGH group, are counted by enumeration.  The header's
// functions must give back the word from its LOW group and LOW rank, and its
// HIGH mask and HIGH rank from the word.
//
// `exchange_host_check big n low samples seed`: words of width n+1 (and n) at
// seeded random global ranks; the LOW rank of each is summed with its LOW
// group's packed table (cell::build_group_table / pack_group_quads, the tables
// the group sweep uses), decoded back, and its HIGH rank from the header's
// group-row table must equal the one summed with its HIGH group's own table
// (build_group_table with the HIGH occupancy): two independent routes.
// Prints one line per case and "ALL OK" at the end; exits non-zero on the
// first mismatch.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "../../oneesan-gpu_amd/csrc/ofp_exchange_impl.h"

extern "C" {
uint64_t or_unrank(uint64_t r, int width);
uint64_t or_main_count(int n);
uint64_t or_blocked_count(int n);
}

using namespace ofp;

static uint64_t g_d[kDpRows][kDpCols];
static cell::Quad<uint32_t> g_plain[kDpRows][kDpCols];

static void plain_tables() {
    for (int w = 0; w < kDpRows; ++w)
        for (int h = 0; h < kDpCols; ++h) g_d[w][h] = 0;
    g_d[0][0] = 1;
    for (int w = 1; w < kDpRows; ++w)
        for (int h = 0; h < kDpCols; ++h)
            g_d[w][h] = g_d[w - 1][h] + (h > 0 ? g_d[w - 1][h - 1] : 0) + (h + 1 < kDpCols ? g_d[w - 1][h + 1] : 0);
    cell::pack_quads<uint32_t>(g_d, g_plain);
}

static uint32_t occ(uint64_t word, int lo, int n) {
    uint32_t m = 0;
    for (int i = 0; i < n; ++i)
        if ((word >> (2 * (lo + i))) & 3u) m |= 1u << i;
    return m;
}

// Rank of a word (positions top..0) with a group's completion table f and its
// occupancy o: the sum of group_entry contributions.
static uint64_t group_rank(uint64_t word, int top, const cell::Occupancy& o, const uint64_t (*f)[kDpCols]) {
    uint64_t r = 0;
    int h = 1;
    for (int pos = top; pos >= 0; --pos) {
        const int v = (int)((word >> (2 * pos)) & 3u);
        uint64_t zn, zc;
        cell::group_entry(o, f, pos, h, zn, zc);
        r += v == 2 ? zc : (v == 1 ? zn : 0);
        h += v == 2 ? 1 : (v == 1 ? -1 : 0);
    }
    return r;
}

// LOW group packed table of mask m (both numberings fix positions 0..low-1).
static void low_table(int W, int low, uint32_t m, cell::Quad<uint32_t> (*q)[kDpCols]) {
    cell::Occupancy om, ob;
    cell::group_occupancy(false, W, low, m, om, ob);
    static uint64_t fm[kDpRows][kDpCols], fb[kDpRows][kDpCols];
    cell::build_group_table(om, fm);
    cell::build_group_table(ob, fb);
    cell::pack_group_quads<uint32_t>(ob, fb, om, fm, q);
}

// The block route (decode_top + top_high + the group's bottom table) must give
// the rank decode_rank gives; the bottom table holds f[base][h] entries per h.
static bool block_route_ok(const cell::Quad<uint32_t> (*q)[kDpCols], bool mainw, int top, int base, uint32_t lrank,
                           const uint32_t* ht, const std::vector<uint32_t>& lut, const uint32_t* off,
                           const uint64_t (*f)[kDpCols], uint32_t want_mask, uint32_t want_rank) {
    const xchg::Top t = mainw ? xchg::decode_top<true>(q, top, base, lrank) : xchg::decode_top<false>(q, top, base, lrank);
    if (t.h < 0 || t.h >= kDpCols) return false;
    if ((uint64_t)(off[t.h + 1] - off[t.h]) != f[base][t.h] || t.r_rest >= off[t.h + 1] - off[t.h]) return false;
    const uint32_t r = xchg::top_high(t, ht) + lut[off[t.h] + t.r_rest];
    return t.mask == want_mask && r == want_rank;
}

static int small(int nmax) {
    int cases = 0;
    static cell::Quad<uint32_t> q[kDpRows][kDpCols];
    for (int n = 2; n <= nmax; ++n) {
        const int W = n + 1;
        for (int low = 1; low <= W - 2; ++low) {
            const int nf = W - 1 - low;
            std::vector<uint32_t> htm((size_t)((1u << nf) - 1u) * kDpCols), htb(htm.size());
            xchg::build_high_rows(g_d, low + 1, nf, htm.data());
            xchg::build_high_rows(g_d, low, nf, htb.data());
            for (int kind = 0; kind < 2; ++kind) {
                const bool mainw = kind == 0;
                const int width = mainw ? W : W - 1, top = width - 1, base = mainw ? low + 1 : low;
                const uint64_t cnt = mainw ? or_main_count(n) : or_blocked_count(n);
                std::vector<uint64_t> lowc(1u << low, 0), highc(1u << nf, 0);
                std::vector<uint64_t> words(cnt);
                for (uint64_t i = 0; i < cnt; ++i) words[i] = or_unrank(i, width);
                // LOW rank = count of earlier words of the same LOW group; likewise HIGH
                std::vector<uint64_t> lrank(cnt), hrank(cnt);
                for (uint64_t i = 0; i < cnt; ++i) {
                    lrank[i] = lowc[occ(words[i], 0, low)]++;
                    hrank[i] = highc[occ(words[i], base, nf)]++;
                }
                uint32_t cur = ~0u;
                static uint64_t lf[kDpRows][kDpCols];
                std::vector<uint32_t> lut;
                uint32_t off[kDpCols + 1];
                for (uint64_t i = 0; i < cnt; ++i) {
                    const uint64_t w = words[i];
                    const uint32_t lm = occ(w, 0, low);
                    if (lm != cur) {
                        low_table(W, low, lm, q);
                        cell::Occupancy om, ob;
                        cell::group_occupancy(false, W, low, lm, om, ob);
                        cell::build_group_table(om, lf);
                        xchg::build_bottom_lut(om, lf, g_d, base, lut, off);
                        cur = lm;
                    }
                    const uint64_t got = mainw ? xchg::decode_group_word<true>(q, top, 0, (uint32_t)lrank[i])
                                               : xchg::decode_group_word<false>(q, top, 0, (uint32_t)lrank[i]);
                    const uint64_t top_only = mainw ? xchg::decode_group_word<true>(q, top, base, (uint32_t)lrank[i])
                                                    : xchg::decode_group_word<false>(q, top, base, (uint32_t)lrank[i]);
                    const uint32_t hm = xchg::high_mask(w, base, nf);
                    const uint32_t hr =
                        xchg::high_rank(w, hm, top, base, mainw ? htm.data() : htb.data(), g_plain);
                    uint64_t fw = 0;
                    uint32_t fm = 0;
                    const uint32_t fr =
                        mainw ? xchg::decode_rank<true>(q, g_plain, htm.data(), top, base, (uint32_t)lrank[i], fw, fm)
                              : xchg::decode_rank<false>(q, g_plain, htb.data(), top, base, (uint32_t)lrank[i], fw, fm);
                    const bool blk_ok = block_route_ok(q, mainw, top, base, (uint32_t)lrank[i],
                                                       mainw ? htm.data() : htb.data(), lut, off, lf, hm, hr);
                    if (got != w || xchg::high_mask(top_only, base, nf) != hm || hm != occ(w, base, nf) ||
                        hr != hrank[i] || fw != w || fm != hm || fr != hr || !blk_ok) {
                        printf("MISMATCH n=%d low=%d kind=%s rank=%llu word=%llx decoded=%llx high mask %u/%u rank "
                               "%u/%llu\n",
                               n, low, mainw ? "main" : "blocked", (unsigned long long)i, (unsigned long long)w,
                               (unsigned long long)got, hm, occ(w, base, nf), hr, (unsigned long long)hrank[i]);
                        return 1;
                    }
                }
                // lowc/highc must partition the state (every group counted)
                ++cases;
            }
        }
        printf("ok n=%d all splits, main and blocked words\n", n);
    }
    printf("cases %d\n", cases);
    return 0;
}

static int big(int n, int low, int samples, unsigned seed) {
    const int W = n + 1, nf = W - 1 - low;
    std::vector<uint32_t> htm((size_t)((1u << nf) - 1u) * kDpCols), htb(htm.size());
    xchg::build_high_rows(g_d, low + 1, nf, htm.data());
    xchg::build_high_rows(g_d, low, nf, htb.data());
    std::mt19937_64 rng(seed);
    static cell::Quad<uint32_t> q[kDpRows][kDpCols];
    static uint64_t f[kDpRows][kDpCols];
    for (int kind = 0; kind < 2; ++kind) {
        const bool mainw = kind == 0;
        const int width = mainw ? W : W - 1, top = width - 1, base = mainw ? low + 1 : low;
        const uint64_t cnt = mainw ? or_main_count(n) : or_blocked_count(n);
        for (int s = 0; s < samples; ++s) {
            const uint64_t w = or_unrank(rng() % cnt, width);
            // LOW rank through the LOW group's own table, decoded back with the packed table
            const uint32_t lm = occ(w, 0, low);
            cell::Occupancy om, ob;
            cell::group_occupancy(false, W, low, lm, om, ob);
            cell::build_group_table(mainw ? om : ob, f);
            const uint64_t lr = group_rank(w, top, mainw ? om : ob, f);
            low_table(W, low, lm, q);
            static uint64_t lf[kDpRows][kDpCols];
            std::vector<uint32_t> lut;
            uint32_t off[kDpCols + 1];
            cell::build_group_table(om, lf);
            xchg::build_bottom_lut(om, lf, g_d, base, lut, off);
            const uint64_t got = mainw ? xchg::decode_group_word<true>(q, top, 0, (uint32_t)lr)
                                       : xchg::decode_group_word<false>(q, top, 0, (uint32_t)lr);
            // HIGH rank through the HIGH group's own table
            const uint32_t hm = xchg::high_mask(w, base, nf);
            cell::group_occupancy(true, W, low, hm, om, ob);
            cell::build_group_table(mainw ? om : ob, f);
            const uint64_t want = group_rank(w, top, mainw ? om : ob, f);
            const uint32_t hr = xchg::high_rank(w, hm, top, base, mainw ? htm.data() : htb.data(), g_plain);
            uint64_t fw = 0;
            uint32_t fm = 0;
            const uint32_t fr = mainw ? xchg::decode_rank<true>(q, g_plain, htm.data(), top, base, (uint32_t)lr, fw, fm)
                                      : xchg::decode_rank<false>(q, g_plain, htb.data(), top, base, (uint32_t)lr, fw, fm);
            const bool blk_ok = block_route_ok(q, mainw, top, base, (uint32_t)lr, mainw ? htm.data() : htb.data(),
                                               lut, off, lf, hm, hr);
            if (lr >> 32 || got != w || want >> 32 || hr != (uint32_t)want || fw != w || fm != hm || fr != hr ||
                !blk_ok) {
                printf("MISMATCH big n=%d low=%d kind=%s word=%llx low rank %llu decoded %llx high rank %u/%llu\n",
                       n, low, mainw ? "main" : "blocked", (unsigned long long)w, (unsigned long long)lr,
                       (unsigned long long)got, hr, (unsigned long long)want);
                return 1;
            }
        }
        printf("ok big n=%d low=%d %s: %d words\n", n, low, mainw ? "main" : "blocked", samples);
    }
    return 0;
}

int main(int argc, char** argv) {
    plain_tables();
    if (argc >= 3 && std::string(argv[1]) == "small") {
        if (small(atoi(argv[2]))) return 1;
    } else if (argc >= 6 && std::string(argv[1]) == "big") {
        if (big(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), (unsigned)atoi(argv[5]))) return 1;
    } else {
        fprintf(stderr, "usage: exchange_host_check small NMAX | big n low samples seed\n");
        return 2;
    }
    printf("ALL OK\n");
    return 0;
}
