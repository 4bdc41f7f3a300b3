// Host check of the occupancy-group placement (oneesan-gpu_amd/csrc/
// ofp_cell_impl.h: group_occupancy, build_group_table, pack_group_quads): the
// cell body, run on the CPU with each group's packed table over the group's
// words alone (ranks local to the group, arrays holding only the group), for
// the high cells p = W-1..low+1 of a row over LOW groups and for the low cells
// p = low..1 over HIGH groups (with blocked input), against the CPU oracle's
// cells on the whole state (oracle/gridfp_oracle.c or_apply_cells) on seeded
// random vectors.
//
// Also checked: the group counts F[W][1], F[W-1][1] against enumeration, and
// that no access of the body leaves the group (every load and store index is
// range-checked).
//
// `group_host_check big|bighigh n low mask...` checks single LOW or HIGH groups at full size (n up
// to 27, where the whole state is out of reach): the group's words are
// enumerated depth-first in lexicographic order (their global ranks, from the
// oracle, must come out increasing, so the enumeration index is the rank in
// the group), and the oracle's forward transition (or_include /
// or_blocked_exclude, the push form of one cell) is applied to the group's
// words alone, every target looked up in the group; the body's result must
// equal it.  Built and run by tests/test_group.py; prints one line per case
// and "ALL OK" at the end, exits non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <unordered_map>
#include <vector>

#include "../../oneesan-gpu_amd/csrc/ofp_cell_impl.h"
#include "../../oneesan-gpu_amd/csrc/ofp_pair_impl.h"

extern "C" {
int or_include(uint64_t m, int width, int p, uint64_t* out, int* blocked);
uint64_t or_blocked_exclude(uint64_t b, int p);
uint64_t or_rank(uint64_t m, int width);
int or_apply_cells(int n, uint32_t mod, uint32_t* main, uint32_t* blk, int p_hi, int p_lo, int blk_valid,
                   int use_pull);
uint64_t or_main_count(int n);
uint64_t or_blocked_count(int n);
uint64_t or_unrank(uint64_t r, int width);
}

using namespace ofp;

static bool g_oob = false;

// State-major arrays of one group with stride K; every index is range-checked.
template <int KC>
struct Checked {
    uint32_t* main;
    uint32_t* blk;
    uint64_t nm, nb;
    int K, off;
    void lm(uint64_t i, cell::Vals<KC>& v) const {
        if (i >= nm) { g_oob = true; return; }
        cell::load_vals<KC, cell::kAoS>(main, i, K, off, v);
    }
    void sm(uint64_t i, const cell::Vals<KC>& v) const {
        if (i >= nm) { g_oob = true; return; }
        cell::store_vals<KC, cell::kAoS>(main, i, K, off, v);
    }
    void lb(uint64_t b, cell::Vals<KC>& v) const {
        if (b >= nb) { g_oob = true; return; }
        cell::load_vals<KC, cell::kAoS>(blk, b, K, off, v);
    }
    void sb(uint64_t b, const cell::Vals<KC>& v) const {
        if (b >= nb) { g_oob = true; return; }
        cell::store_vals<KC, cell::kAoS>(blk, b, K, off, v);
    }
};

// The group of a word under occupancy constraint o (its bits at o's positions).
static uint32_t group_of(uint64_t word, const cell::Occupancy& o) {
    uint32_t m = 0;
    for (int i = 0; i < o.n; ++i)
        if ((word >> (2 * (o.lo + i))) & 3u) m |= 1u << i;
    return m;
}

// One pair pass (cells p, p-1) over one group's arrays, in place, launches in
// the kernels' order d = -1, 0, 1, -2, 2, classes in rank order or reversed.
template <int KC, int D>
static void pair_d(const pair::PairLaunch& pa, const cell::Tab<uint32_t>& T, const cell::Occupancy& om,
                   const uint64_t (*fm)[kDpCols], const pair::Io<KC, uint32_t>& io, bool rev) {
    static cell::Pair<uint32_t> ct[kDpRows][kDpCols];
    pair::pack_group_class_table<uint32_t, kDpCols>(om, fm, pa.p, D, ct);
    const uint64_t cnt = pair::group_class_count(om, fm, pa.W, pa.p, D);
    for (uint64_t t = 0; t < cnt; ++t) {
        const uint64_t c = rev ? cnt - 1 - t : t;
        const pair::Class<uint32_t> k = pair::decode_class<uint32_t>(T, ct, pa.W, pa.p, c);
        pair::pair_body<KC, uint32_t, D, 0>(pa, T, k, io);
    }
}

// Runs the cells of one half (high = false: W-1..low+1 over LOW groups, true:
// low..1 over HIGH groups) group by group and compares with the oracle.  Pairs:
// cells p, p-1 as pair passes where both are in the half (the group's class
// tables, pack_group_class_table), the last odd cell alone; Rev: the classes
// of each launch in reverse order.
template <int V, bool Pairs = false, bool Rev = false>
static int check(int n, int low, bool high, const uint32_t* mods, unsigned seed) {
    constexpr int KC = 3, K = 4;
    const int W = n + 1;
    const int fixed = high ? W - 1 - low : low, p_hi = high ? low : W - 1, p_lo = high ? 1 : low + 1;
    const uint64_t nm = or_main_count(n), nb = or_blocked_count(n);
    cell::Occupancy om0, ob0;
    cell::group_occupancy(high, W, low, 0, om0, ob0);
    // group of every word and its rank inside the group (global order)
    std::vector<uint32_t> gm(nm), gb(nb);
    std::vector<uint64_t> lm(nm), lb(nb);
    std::vector<uint64_t> cm(1u << fixed, 0), cb(1u << fixed, 0);
    for (uint64_t i = 0; i < nm; ++i) {
        gm[i] = group_of(or_unrank(i, W), om0);
        lm[i] = cm[gm[i]]++;
    }
    for (uint64_t i = 0; i < nb; ++i) {
        gb[i] = group_of(or_unrank(i, W - 1), ob0);
        lb[i] = cb[gb[i]]++;
    }
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> mi(nm * KC), bi((nb ? nb : 1) * KC, 0);
    for (uint64_t i = 0; i < nm; ++i)
        for (int j = 0; j < KC; ++j) mi[i * KC + j] = (uint32_t)(rng() % mods[j]);
    if (high)
        for (uint64_t i = 0; i < nb; ++i)
            for (int j = 0; j < KC; ++j) bi[i * KC + j] = (uint32_t)(rng() % mods[j]);
    std::vector<uint32_t> mo(nm * KC, 0), bo((nb ? nb : 1) * KC, 0);
    // per group: gather, run the cells with the group's table, scatter back
    static uint64_t fm[kDpRows][kDpCols], fb[kDpRows][kDpCols];
    static cell::Quad<uint32_t> q[kDpRows][kDpCols];
    int groups = 0;
    for (uint32_t g = 0; g < (1u << fixed); ++g) {
        cell::Occupancy om, ob;
        cell::group_occupancy(high, W, low, g, om, ob);
        cell::build_group_table(om, fm);
        cell::build_group_table(ob, fb);
        if (fm[W][1] != cm[g] || fb[W - 1][1] != cb[g]) {
            printf("COUNT n=%d low=%d high=%d group=%u main %llu/%llu blocked %llu/%llu\n", n, low, high, g,
                   (unsigned long long)fm[W][1], (unsigned long long)cm[g], (unsigned long long)fb[W - 1][1],
                   (unsigned long long)cb[g]);
            return 1;
        }
        if (!cm[g]) continue;
        ++groups;
        cell::pack_group_quads<uint32_t>(ob, fb, om, fm, q);
        cell::Tab<uint32_t> T{q};
        // kGuard records past each array hold a sentinel: a read there spoils a
        // residue, a write there is caught below
        constexpr uint64_t kGuard = 64;
        constexpr uint32_t kSentinel = 0xA5A5A5A5u;
        std::vector<uint32_t> am((cm[g] + kGuard) * K, 0), ab(((cb[g] ? cb[g] : 1) + kGuard) * K, 0);
        for (uint64_t i = cm[g] * K; i < am.size(); ++i) am[i] = kSentinel;
        for (uint64_t i = (cb[g] ? cb[g] : 1) * K; i < ab.size(); ++i) ab[i] = kSentinel;
        for (uint64_t i = 0; i < nm; ++i)
            if (gm[i] == g)
                for (int j = 0; j < KC; ++j) am[lm[i] * K + j] = mi[i * KC + j];
        for (uint64_t i = 0; i < nb; ++i)
            if (gb[i] == g)
                for (int j = 0; j < KC; ++j) ab[lb[i] * K + j] = bi[i * KC + j];
        Checked<KC> mem{am.data(), ab.data(), cm[g], cb[g], K, 0};
        CellLaunch a{};
        a.W = W;
        a.K = K;
        a.off = 0;
        a.kc = KC;
        a.layout = 1;
        for (int j = 0; j < KC; ++j) a.mod[j] = mods[j];
        for (int p = p_hi; p >= p_lo;) {
            if (Pairs && p - 1 >= p_lo && p >= 2) {
                pair::PairLaunch pa{};
                pa.main_in = am.data();
                pa.blk_in = ab.data();
                pa.main_out = am.data();
                pa.blk_out = ab.data();
                pa.W = W;
                pa.p = p;
                pa.read_blk = p == W - 1 ? 0 : 1;
                pa.K = K;
                pa.off = 0;
                pa.kc = KC;
                for (int j = 0; j < KC; ++j) pa.mod[j] = mods[j];
                const pair::Io<KC, uint32_t> io{am.data(), ab.data(), am.data(), ab.data(), K, 0};
                pair_d<KC, -1>(pa, T, om, fm, io, Rev);
                pair_d<KC, 0>(pa, T, om, fm, io, Rev);
                pair_d<KC, 1>(pa, T, om, fm, io, Rev);
                pair_d<KC, -2>(pa, T, om, fm, io, Rev);
                pair_d<KC, 2>(pa, T, om, fm, io, Rev);
                bool spoiled = false;
                for (uint64_t i = cm[g] * K; i < am.size(); ++i) spoiled |= am[i] != kSentinel;
                for (uint64_t i = (cb[g] ? cb[g] : 1) * K; i < ab.size(); ++i) spoiled |= ab[i] != kSentinel;
                if (spoiled) {
                    printf("OUT OF GROUP (pair) n=%d low=%d high=%d group=%u p=%d\n", n, low, high, g, p);
                    return 1;
                }
                p -= 2;
                continue;
            }
            a.p = p;
            a.read_blk = p == W - 1 ? 0 : 1;
            for (uint64_t b = 0; b < cb[g]; ++b) cell::cell_one<KC, uint32_t, cell::kAoS, V>(a, T, b, mem);
            if (g_oob) {
                printf("OUT OF GROUP n=%d low=%d high=%d group=%u p=%d\n", n, low, high, g, p);
                return 1;
            }
            --p;
        }
        for (uint64_t i = 0; i < nm; ++i)
            if (gm[i] == g)
                for (int j = 0; j < KC; ++j) mo[i * KC + j] = am[lm[i] * K + j];
        for (uint64_t i = 0; i < nb; ++i)
            if (gb[i] == g)
                for (int j = 0; j < KC; ++j) bo[i * KC + j] = ab[lb[i] * K + j];
    }
    for (int j = 0; j < KC; ++j) {
        std::vector<uint32_t> em(nm), eb(nb ? nb : 1, 0);
        for (uint64_t i = 0; i < nm; ++i) em[i] = mi[i * KC + j];
        for (uint64_t i = 0; i < nb; ++i) eb[i] = bi[i * KC + j];
        if (or_apply_cells(n, mods[j], em.data(), eb.data(), p_hi, p_lo, high ? 1 : 0, 1) != 0) return 2;
        for (uint64_t i = 0; i < nm; ++i)
            if (mo[i * KC + j] != em[i]) {
                printf("MISMATCH main n=%d low=%d high=%d V=%d j=%d i=%llu got=%u want=%u\n", n, low, high, V, j,
                       (unsigned long long)i, mo[i * KC + j], em[i]);
                return 1;
            }
        if (!high)
            for (uint64_t i = 0; i < nb; ++i)
                if (bo[i * KC + j] != eb[i]) {
                    printf("MISMATCH blocked n=%d low=%d V=%d j=%d i=%llu got=%u want=%u\n", n, low, V, j,
                           (unsigned long long)i, bo[i * KC + j], eb[i]);
                    return 1;
                }
    }
    printf("ok n=%d low=%d %s V=%d%s groups=%d cells=%d..%d\n", n, low, high ? "high" : "low", V,
           Pairs ? (Rev ? " pairs reversed" : " pairs") : "", groups, p_hi, p_lo);
    return 0;
}

// Words of `width` positions in the group, in lexicographic order (N < R < L
// from the top, entered at height 1, ending at 0).
static void group_words(int width, const cell::Occupancy& o, const uint64_t (*f)[kDpCols], std::vector<uint64_t>& out) {
    std::function<void(int, int, uint64_t)> rec = [&](int pos, int h, uint64_t word) {
        if (pos < 0) {
            if (h == 0) out.push_back(word);
            return;
        }
        const int fix = o.at(pos);
        for (int v = 0; v < 3; ++v) {
            if (fix >= 0 && (v != 0) != (fix == 1)) continue;
            const int nh = h + (v == 2) - (v == 1);
            if (nh < 0 || nh >= kDpCols || f[pos][nh] == 0) continue;
            rec(pos - 1, nh, word | ((uint64_t)v << (2 * pos)));
        }
    };
    rec(width - 1, 1, 0);
}

static int check_big(int n, int low, bool high, uint32_t g, const uint32_t* mods, unsigned seed, bool pairs) {
    constexpr int KC = 3, K = 4;
    const int W = n + 1;
    const int p_hi = high ? low : W - 1, p_lo = high ? 1 : low + 1;
    static uint64_t fm[kDpRows][kDpCols], fb[kDpRows][kDpCols];
    static cell::Quad<uint32_t> q[kDpRows][kDpCols];
    cell::Occupancy om, ob;
    cell::group_occupancy(high, W, low, g, om, ob);
    cell::build_group_table(om, fm);
    cell::build_group_table(ob, fb);
    std::vector<uint64_t> wm, wb;
    group_words(W, om, fm, wm);
    group_words(W - 1, ob, fb, wb);
    if (wm.size() != fm[W][1] || wb.size() != fb[W - 1][1]) {
        printf("COUNT big n=%d low=%d group=%u\n", n, low, g);
        return 1;
    }
    std::unordered_map<uint64_t, uint64_t> im, ib;
    for (uint64_t i = 0; i < wm.size(); ++i) {
        if (i && or_rank(wm[i], W) <= or_rank(wm[i - 1], W)) {
            printf("ORDER big main n=%d group=%u i=%llu\n", n, g, (unsigned long long)i);
            return 1;
        }
        im[wm[i]] = i;
    }
    for (uint64_t i = 0; i < wb.size(); ++i) {
        if (i && or_rank(wb[i], W - 1) <= or_rank(wb[i - 1], W - 1)) {
            printf("ORDER big blocked n=%d group=%u i=%llu\n", n, g, (unsigned long long)i);
            return 1;
        }
        ib[wb[i]] = i;
    }
    const uint64_t nm = wm.size(), nb = wb.size();
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> mi(nm * K, 0), bi((nb ? nb : 1) * K, 0);
    for (uint64_t i = 0; i < nm; ++i)
        for (int j = 0; j < KC; ++j) mi[i * K + j] = (uint32_t)(rng() % mods[j]);
    if (high)
        for (uint64_t i = 0; i < nb; ++i)
            for (int j = 0; j < KC; ++j) bi[i * K + j] = (uint32_t)(rng() % mods[j]);
    std::vector<uint32_t> em = mi, eb = bi;  // the oracle's run
    cell::pack_group_quads<uint32_t>(ob, fb, om, fm, q);
    cell::Tab<uint32_t> T{q};
    Checked<KC> mem{mi.data(), bi.data(), nm, nb, K, 0};
    CellLaunch a{};
    a.W = W;
    a.K = K;
    a.kc = KC;
    a.layout = 1;
    for (int j = 0; j < KC; ++j) a.mod[j] = mods[j];
    for (int p = p_hi; p >= p_lo;) {
        const bool pr = pairs && p - 1 >= p_lo && p >= 2;
        if (pr) {
            pair::PairLaunch pa{};
            pa.main_in = mi.data();
            pa.blk_in = bi.data();
            pa.main_out = mi.data();
            pa.blk_out = bi.data();
            pa.W = W;
            pa.p = p;
            pa.read_blk = p == W - 1 ? 0 : 1;
            pa.K = K;
            pa.off = 0;
            pa.kc = KC;
            for (int j = 0; j < KC; ++j) pa.mod[j] = mods[j];
            const pair::Io<KC, uint32_t> io{mi.data(), bi.data(), mi.data(), bi.data(), K, 0};
            pair_d<KC, -1>(pa, T, om, fm, io, false);
            pair_d<KC, 0>(pa, T, om, fm, io, false);
            pair_d<KC, 1>(pa, T, om, fm, io, false);
            pair_d<KC, -2>(pa, T, om, fm, io, false);
            pair_d<KC, 2>(pa, T, om, fm, io, false);
        } else {
            a.p = p;
            a.read_blk = p == W - 1 ? 0 : 1;
            for (uint64_t b = 0; b < nb; ++b) cell::cell_one<KC, uint32_t, cell::kAoS, 2>(a, T, b, mem);
            if (g_oob) {
                printf("OUT OF GROUP big n=%d low=%d group=%u p=%d\n", n, low, g, p);
                return 1;
            }
        }
        // oracle: push form of each cell over the group's words
        const int q_lo = pr ? p - 1 : p;
        for (int q = p; q >= q_lo; --q) {
            std::vector<uint32_t> mo = em, bo(eb.size(), 0);
            for (uint64_t i = 0; i < nm; ++i) {
                uint64_t t;
                int blocked;
                if (!or_include(wm[i], W, q, &t, &blocked)) continue;
                auto& map = blocked ? ib : im;
                auto it = map.find(t);
                if (it == map.end()) {
                    printf("TRANSITION LEAVES GROUP n=%d group=%u p=%d\n", n, g, q);
                    return 1;
                }
                uint32_t* dst = blocked ? &bo[it->second * K] : &mo[it->second * K];
                for (int j = 0; j < KC; ++j) dst[j] = cell::addm(dst[j], em[i * K + j], mods[j]);
            }
            if (q != W - 1)
                for (uint64_t i = 0; i < nb; ++i) {
                    auto it = im.find(or_blocked_exclude(wb[i], q));
                    if (it == im.end()) {
                        printf("EXCLUDE LEAVES GROUP n=%d group=%u p=%d\n", n, g, q);
                        return 1;
                    }
                    for (int j = 0; j < KC; ++j)
                        mo[it->second * K + j] = cell::addm(mo[it->second * K + j], eb[i * K + j], mods[j]);
                }
            em.swap(mo);
            eb.swap(bo);
        }
        for (uint64_t i = 0; i < nm; ++i)
            for (int j = 0; j < KC; ++j)
                if (mi[i * K + j] != em[i * K + j]) {
                    printf("MISMATCH big main n=%d group=%u p=%d i=%llu j=%d\n", n, g, p, (unsigned long long)i, j);
                    return 1;
                }
        if (q_lo >= 2)
            for (uint64_t i = 0; i < nb; ++i)
                for (int j = 0; j < KC; ++j)
                    if (bi[i * K + j] != eb[i * K + j]) {
                        printf("MISMATCH big blocked n=%d group=%u p=%d i=%llu j=%d\n", n, g, p,
                               (unsigned long long)i, j);
                        return 1;
                    }
        p = q_lo - 1;
    }
    printf("ok big n=%d low=%d %s%s group=%u main=%llu blocked=%llu cells=%d..%d\n", n, low, high ? "high" : "low",
           pairs ? " pairs" : "", g, (unsigned long long)nm, (unsigned long long)nb, p_hi, p_lo);
    return 0;
}

int main(int argc, char** argv) {
    const uint32_t bmods[3] = {4294967291u, 2u, 4294966997u};
    if (argc > 3 && (!strncmp(argv[1], "big", 3))) {
        const bool high = !strncmp(argv[1], "bighigh", 7);
        const bool pairs = strstr(argv[1], "pair") != nullptr;  // bigpair, bighighpair: pair passes
        const int n = atoi(argv[2]), low = atoi(argv[3]);
        for (int i = 4; i < argc; ++i)
            if (check_big(n, low, high, (uint32_t)strtoul(argv[i], nullptr, 10), bmods, 77u + i, pairs)) return 1;
        printf("ALL OK\n");
        return 0;
    }
    const int nmax = argc > 1 ? atoi(argv[1]) : 10;
    const uint32_t mods[3] = {4294967291u, 2u, 4294966997u};
    for (int n = 2; n <= nmax; ++n) {
        const int W = n + 1;
        for (int low = 1; low <= W - 2; ++low)
            for (int high = 0; high < 2; ++high) {
                if (check<0>(n, low, high, mods, 100u * n + low + 50 * high)) return 1;
                if (check<2>(n, low, high, mods, 200u * n + low + 50 * high)) return 1;
                if (check<2, true, false>(n, low, high, mods, 500u * n + low + 50 * high)) return 1;
                if (check<2, true, true>(n, low, high, mods, 600u * n + low + 50 * high)) return 1;
                if (n % 2 == 0 && (check<1>(n, low, high, mods, 300u * n + low + 50 * high) ||
                                   check<3>(n, low, high, mods, 400u * n + low + 50 * high)))
                    return 1;
            }
    }
    printf("ALL OK\n");
    return 0;
}
