// Host check of the pair pass (oneesan-gpu_amd/csrc/ofp_pair_impl.h): the
// per-class body, run on the CPU over every class of (p, d) IN PLACE (one
// array, launches in the kernels' order d = -1, 0, 1, -3, -2, 2, 3), against
// the CPU oracle's two-cell window (oracle/gridfp_oracle.c or_apply_cells) on
// seeded random vectors.  Inside each launch the classes are taken in rank
// order and, in a second run, in reverse, so that every foreign read of a
// class's word happens once before and once after that class writes it.
// Built and run by tests/test_pair_host.py; prints one line per case and
// "ALL OK" at the end, exits non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "../../oneesan-gpu_amd/csrc/ofp_pair_impl.h"

extern "C" int or_apply_cells(int n, uint32_t mod, uint32_t* main, uint32_t* blk, int p_hi, int p_lo,
                              int blk_valid, int use_pull);

using namespace ofp;

static uint64_t D0[kDpRows][kDpCols];

static void build_d() {
    memset(D0, 0, sizeof D0);
    D0[0][0] = 1;
    for (int w = 1; w < kDpRows; ++w)
        for (int h = 0; h < kDpCols; ++h)
            D0[w][h] = D0[w - 1][h] + (h > 0 ? D0[w - 1][h - 1] : 0) + (h + 1 < kDpCols ? D0[w - 1][h + 1] : 0);
}

template <int KC, typename Idx, int D, int GF>
static void run_d(const pair::PairLaunch& a, const cell::Tab<Idx>& T, const pair::Io<KC, Idx>& io, bool rev) {
    constexpr int C = cell::LdsCols<Idx>::value;
    static cell::Pair<Idx> ct[kDpRows][C];
    pair::pack_class_table<Idx, C>(D0, a.p, D, ct);
    const uint64_t n = pair::class_count(D0, a.W, a.p, D);
    for (uint64_t t = 0; t < n; ++t) {
        const uint64_t c = rev ? n - 1 - t : t;
        const pair::Class<Idx> k = pair::decode_class<Idx>(T, ct, a.W, a.p, c);
        pair::pair_body<KC, Idx, D, GF>(a, T, k, io);
    }
}

template <int KC, typename Idx, int GF>
static int check(int n, int p, bool read_blk, const uint32_t* mods, unsigned seed, bool rev) {
    const int W = n + 1;
    const uint64_t nm = D0[W][1], nb = D0[W - 1][1];
    constexpr int C = cell::LdsCols<Idx>::value;
    static cell::Quad<Idx> qfull[kDpRows][kDpCols];
    static cell::Quad<Idx> q[kDpRows][C];
    cell::pack_quads<Idx>(D0, qfull);
    for (int r = 0; r < kDpRows; ++r)
        for (int h = 0; h < C; ++h) q[r][h] = qfull[r][h];
    cell::Tab<Idx> T{q};
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> mi(nm * KC), bi((nb ? nb : 1) * KC);
    for (uint64_t i = 0; i < nm; ++i)
        for (int j = 0; j < KC; ++j) mi[i * KC + j] = (uint32_t)(rng() % mods[j]);
    for (uint64_t i = 0; i < nb; ++i)
        for (int j = 0; j < KC; ++j) bi[i * KC + j] = (uint32_t)(rng() % mods[j]);
    pair::PairLaunch a{};
    a.W = W;
    a.p = p;
    a.read_blk = read_blk ? 1 : 0;
    a.K = KC;
    a.off = 0;
    a.kc = KC;
    for (int j = 0; j < KC; ++j) a.mod[j] = mods[j];
    const std::vector<uint32_t> m0 = mi, b0 = bi;  // the inputs, for the oracle
    std::vector<uint32_t>& mo = mi;
    std::vector<uint32_t>& bo = bi;
    pair::Io<KC, Idx> io{mi.data(), bi.data(), mi.data(), bi.data(), KC, 0};
    run_d<KC, Idx, -1, GF>(a, T, io, rev);
    run_d<KC, Idx, 0, GF>(a, T, io, rev);
    run_d<KC, Idx, 1, GF>(a, T, io, rev);
    run_d<KC, Idx, -3, GF>(a, T, io, rev);
    run_d<KC, Idx, -2, GF>(a, T, io, rev);
    run_d<KC, Idx, 2, GF>(a, T, io, rev);
    run_d<KC, Idx, 3, GF>(a, T, io, rev);
    for (int j = 0; j < KC; ++j) {
        std::vector<uint32_t> em(nm), eb(nb ? nb : 1);
        for (uint64_t i = 0; i < nm; ++i) em[i] = m0[i * KC + j];
        for (uint64_t i = 0; i < nb; ++i) eb[i] = b0[i * KC + j];
        if (or_apply_cells(n, mods[j], em.data(), eb.data(), p, p - 1, read_blk ? 1 : 0, 1) != 0) return 2;
        for (uint64_t i = 0; i < nm; ++i)
            if (mo[i * KC + j] != em[i]) {
                printf("MISMATCH main n=%d p=%d blk=%d idx=%s j=%d i=%llu got=%u want=%u\n", n, p, read_blk,
                       sizeof(Idx) == 8 ? "u64" : "u32", j, (unsigned long long)i, mo[i * KC + j], em[i]);
                return 1;
            }
        if (p - 1 >= 2)
            for (uint64_t i = 0; i < nb; ++i)
                if (bo[i * KC + j] != eb[i]) {
                    printf("MISMATCH blocked n=%d p=%d blk=%d idx=%s j=%d i=%llu got=%u want=%u\n", n, p, read_blk,
                           sizeof(Idx) == 8 ? "u64" : "u32", j, (unsigned long long)i, bo[i * KC + j], eb[i]);
                    return 1;
                }
    }
    printf("ok n=%d p=%d blk=%d idx=%s k=%d order=%s gathers_first=%d\n", n, p, read_blk,
           sizeof(Idx) == 8 ? "u64" : "u32", KC, rev ? "reverse" : "forward", GF);
    return 0;
}

int main(int argc, char** argv) {
    build_d();
    const int nmax = argc > 1 ? atoi(argv[1]) : 9;
    const uint32_t mods[3] = {4294967291u, 2u, 4294966997u};
    for (int n = 2; n <= nmax; ++n) {
        const int W = n + 1;
        for (int p = W - 1; p >= 2; --p)
            for (int rb = 0; rb < 2; ++rb)
                for (int rev = 0; rev < 2; ++rev) {
                    if (check<3, uint32_t, 0>(n, p, rb, mods, 1000u * n + 10u * p + rb, rev)) return 1;
                    if (check<3, uint32_t, 1>(n, p, rb, mods, 3000u * n + 10u * p + rb, rev)) return 1;
                    if (n % 3 == 0 && check<3, uint64_t, 0>(n, p, rb, mods, 7000u * n + 10u * p + rb, rev)) return 1;
                }
    }
    printf("ALL OK\n");
    return 0;
}
