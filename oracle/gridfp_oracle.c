/*
 * gridfp_oracle.c -- CPU restatement of the reference Grid-FP frontier sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it.
 *
 * What is restated, and from where (reference paths, file:line):
 *   - frontier word: 2 bits per position, N=0 R=1 L=2
 *       src/common/gridfp_transition.hpp:18-33 (mget/mset/msetpair/mshrink/minsert)
 *   - closure matching scans
 *       src/common/gridfp_transition.hpp:66-103
 *   - the included-horizontal-edge transition and the blocked re-expansion
 *       src/common/gridfp_transition.hpp:113-146
 *   - reverse (predecessor) enumeration for blocked targets / p=1 boundary
 *       src/common/gridfp_reverse.hpp:39-98
 *   - Motzkin-height table H_DP and the lexicographic global rank (N<R<L)
 *       src/cuda/b300/oneesan_cuda_gridfp_b300_hbm32_fullmate_dropN_batch.cu:62,402
 *   - one cell step p (forward form) -- literally the reference's CPU oracle
 *       tests/factor_transfer_batch_test.cu:60-72
 *   - the run: main[rank(R<<2(W-1))]=1, W rows x cells p=W-1..1, answer
 *     main[rank(R)]   fullmate_dropN_batch.cu:508-514
 *
 * Two step forms are provided: or_cell_forward() (push, sequential, the
 * literal recurrence) and or_cell_pull() (pull over reverse predecessors,
 * OpenMP-parallel).  Tests pin them against each other on random vectors.
 *
 * Arithmetic: uint32 residues mod p (2 <= p < 2^32), additions only.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_MAXW 28
typedef uint64_t word_t;
enum { VN = 0, VR = 1, VL = 2 };

static uint64_t DP[OR_MAXW + 2][OR_MAXW + 3];
static int dp_ready = 0;

/* H_DP[w][h]: words of w positions, entered at height h, ending at 0, never
 * negative.  fullmate_dropN_batch.cu:62 */
static void build_dp(void) {
    if (dp_ready) return;
    for (int h = 0; h <= OR_MAXW + 2; ++h) DP[0][h] = (h == 0);
    for (int w = 1; w <= OR_MAXW + 1; ++w)
        for (int h = 0; h <= OR_MAXW + 1; ++h) {
            uint64_t x = DP[w - 1][h];
            if (h > 0) x += DP[w - 1][h - 1];
            x += DP[w - 1][h + 1];
            DP[w][h] = x;
        }
    dp_ready = 1;
}

uint64_t or_dp(int w, int h) {
    build_dp();
    if (w < 0 || w > OR_MAXW + 1 || h < 0 || h > OR_MAXW + 1) return 0;
    return DP[w][h];
}
uint64_t or_main_count(int n) { build_dp(); return DP[n + 1][1]; }
uint64_t or_blocked_count(int n) { build_dp(); return DP[n][1]; }

static inline int mget(word_t m, int k) { return (int)((m >> (2 * k)) & 3u); }
static inline int mpair(word_t m, int p) { return (int)((m >> (2 * (p - 1))) & 15u); }
static inline word_t mset(word_t m, int k, int v) {
    word_t z = (word_t)3 << (2 * k);
    return (m & ~z) | ((word_t)v << (2 * k));
}
static inline word_t msetpair(word_t m, int p, int v) {
    word_t z = (word_t)15 << (2 * (p - 1));
    return (m & ~z) | ((word_t)v << (2 * (p - 1)));
}
static inline word_t mshrink(word_t m, int k) {
    word_t mask = ((word_t)1 << (2 * k)) - 1;
    return ((m & ~mask) >> 2) | (m & mask);
}
static inline word_t minsert(word_t m, int k, int v) {
    word_t lowmask = k ? (((word_t)1 << (2 * k)) - 1) : 0;
    return (m & lowmask) | ((word_t)v << (2 * k)) | ((m & ~lowmask) << 2);
}
#define PAIR(a, b) (((a) << 2) | (b))

/* plain scans, gridfp_transition.hpp:66-103 (the non-"FAST" branch) */
static int closure_match_left(word_t t, int p) {
    int q = p - 1, s = 1;
    while (s) {
        --q;
        if (q < 0) return -1;
        int v = mget(t, q);
        if (v == VL) ++s; else if (v == VR) --s;
    }
    return q;
}
static int closure_match_right(word_t t, int width, int p) {
    int q = p, s = 1;
    while (s) {
        ++q;
        if (q >= width) return -1;
        int v = mget(t, q);
        if (v == VL) --s; else if (v == VR) ++s;
    }
    return q;
}

/* gridfp_transition.hpp:113-144.  Returns 1 if valid; *out and *blocked set. */
int or_include(word_t m, int width, int p, word_t* out, int* blocked) {
    word_t t;
    int pr = mpair(m, p), q;
    *blocked = 0;
    switch (pr) {
    case PAIR(VN, VN): *out = msetpair(m, p, PAIR(VL, VR)); return 1;
    case PAIR(VN, VR):
    case PAIR(VN, VL):
        if (p == 1) { *out = msetpair(m, p, pr == PAIR(VN, VR) ? PAIR(VR, VN) : PAIR(VL, VN)); return 1; }
        *out = mshrink(m, p); *blocked = 1; return 1;
    case PAIR(VR, VN): *out = msetpair(m, p, PAIR(VN, VR)); return 1;
    case PAIR(VL, VN): *out = msetpair(m, p, PAIR(VN, VL)); return 1;
    case PAIR(VL, VL):
        t = msetpair(m, p, PAIR(VN, VN));
        q = closure_match_left(t, p);
        if (q < 0) return 0;
        t = mset(t, q, VL);
        if (p == 1) { *out = t; return 1; }
        *out = mshrink(t, p - 1); *blocked = 1; return 1;
    case PAIR(VR, VR):
        t = msetpair(m, p, PAIR(VN, VN));
        q = closure_match_right(t, width, p);
        if (q < 0) return 0;
        t = mset(t, q, VR);
        if (p == 1) { *out = t; return 1; }
        *out = mshrink(t, p - 1); *blocked = 1; return 1;
    case PAIR(VR, VL):
        t = msetpair(m, p, PAIR(VN, VN));
        if (p == 1) { *out = t; return 1; }
        *out = mshrink(t, p - 1); *blocked = 1; return 1;
    default:
        return 0;
    }
}
word_t or_blocked_exclude(word_t b, int p) { return minsert(b, p, VN); }

/* lexicographic rank, N<R<L from the high position, entered at height 1
 * (fullmate_dropN_batch.cu:402) */
uint64_t or_rank(word_t m, int width) {
    build_dp();
    uint64_t r = 0;
    int h = 1;
    for (int pos = width - 1; pos >= 0; --pos) {
        int v = mget(m, pos);
        if (v > VN) r += DP[pos][h];
        if (v > VR && h > 0) r += DP[pos][h - 1];
        if (v == VR) --h; else if (v == VL) ++h;
    }
    return r;
}
word_t or_unrank(uint64_t r, int width) {
    build_dp();
    word_t m = 0;
    int h = 1;
    for (int pos = width - 1; pos >= 0; --pos) {
        uint64_t z = DP[pos][h];
        if (r < z) continue;
        r -= z;
        if (h > 0) {
            z = DP[pos][h - 1];
            if (r < z) { m |= (word_t)VR << (2 * pos); --h; continue; }
            r -= z;
        }
        m |= (word_t)VL << (2 * pos);
        ++h;
    }
    return m;
}
/* legality: entered at height 1, never negative, ends at 0 */
int or_is_legal(word_t m, int width) {
    int h = 1;
    if (width < 32 && (m >> (2 * width)) != 0) return 0;
    for (int pos = width - 1; pos >= 0; --pos) {
        int v = mget(m, pos);
        if (v == 3) return 0;
        if (v == VR) { if (--h < 0) return 0; }
        else if (v == VL) ++h;
    }
    return h == 0;
}

/* gridfp_reverse.hpp:39-83, scalar (non-sparse) form; emits into out[] */
int or_reverse_block_preds(word_t b, int width, int p, word_t* out, int cap) {
    int n = 0;
    int low = mget(b, p - 1);
#define EMIT(x) do { if (n < cap) out[n] = (x); ++n; } while (0)
    if (low == VR || low == VL) { EMIT(minsert(b, p, VN)); return n; }
    if (low != VN) return 0;
    word_t u = minsert(b, p - 1, VN);
    int height = 1;
    for (int q = width - 1; q > p; --q) {
        int v = mget(u, q);
        if (v == VL) ++height; else if (v == VR) --height;
    }
    if (height > 0) EMIT(msetpair(u, p, PAIR(VR, VL)));
    int depth = 1;
    for (int q = p - 2; q >= 0 && depth > 0; --q) {
        int v = mget(u, q);
        if (v == VL && depth == 1) EMIT(mset(msetpair(u, p, PAIR(VL, VL)), q, VR));
        if (v == VL) ++depth; else if (v == VR) --depth;
    }
    depth = 1;
    for (int q = p + 1; q < width && depth > 0; ++q) {
        int v = mget(u, q);
        if (v == VR && depth == 1) EMIT(mset(msetpair(u, p, PAIR(VR, VR)), q, VL));
        if (v == VL) --depth; else if (v == VR) ++depth;
    }
    return n;
}
/* gridfp_reverse.hpp:87-98 */
int or_reverse_boundary_preds(word_t t, int width, word_t* out, int cap) {
    switch (mpair(t, 1)) {
    case PAIR(VL, VR): out[0] = msetpair(t, 1, PAIR(VN, VN)); return 1;
    case PAIR(VN, VR): out[0] = msetpair(t, 1, PAIR(VR, VN)); return 1;
    case PAIR(VN, VL): out[0] = msetpair(t, 1, PAIR(VL, VN)); return 1;
    case PAIR(VR, VN): out[0] = msetpair(t, 1, PAIR(VN, VR)); return 1;
    case PAIR(VL, VN): out[0] = msetpair(t, 1, PAIR(VN, VL)); return 1;
    case PAIR(VN, VN): return or_reverse_block_preds(mshrink(t, 0), width, 1, out, cap);
    default: return 0;
    }
}
#undef EMIT

static inline uint32_t addm(uint32_t a, uint32_t b, uint32_t p) {
    uint64_t s = (uint64_t)a + b;
    return (uint32_t)(s >= p ? s - p : s);
}

/* One cell, forward (push) form: tests/factor_transfer_batch_test.cu:60-72.
 * main_in/blk_in are the vectors before the cell; outputs are written fully.
 * If blk_in is NULL the blocked input is taken as zero. */
void or_cell_forward(int n, int p, uint32_t mod, const uint32_t* main_in,
                     const uint32_t* blk_in, uint32_t* main_out, uint32_t* blk_out) {
    build_dp();
    int W = n + 1;
    uint64_t nm = DP[W][1], nb = DP[W - 1][1];
    memcpy(main_out, main_in, nm * sizeof(uint32_t));
    memset(blk_out, 0, nb * sizeof(uint32_t));
    for (uint64_t i = 0; i < nm; ++i) {
        uint32_t val = main_in[i];
        if (!val) continue;
        word_t m = or_unrank(i, W), t;
        int blocked;
        if (!or_include(m, W, p, &t, &blocked)) continue;
        if (blocked) { uint64_t j = or_rank(t, W - 1); blk_out[j] = addm(blk_out[j], val, mod); }
        else { uint64_t j = or_rank(t, W); main_out[j] = addm(main_out[j], val, mod); }
    }
    if (blk_in)
        for (uint64_t i = 0; i < nb; ++i) {
            uint32_t val = blk_in[i];
            if (!val) continue;
            word_t b = or_unrank(i, W - 1);
            uint64_t j = or_rank(or_blocked_exclude(b, p), W);
            main_out[j] = addm(main_out[j], val, mod);
        }
}

/* Pull form restricted to main targets [t0,t1) and blocked targets [b0,b1)
 * (the full-size arrays are still addressed); used to time a bounded sample. */
void or_cell_pull_range(int n, int p, uint32_t mod, const uint32_t* main_in,
                        const uint32_t* blk_in, uint32_t* main_out, uint32_t* blk_out,
                        int64_t t0, int64_t t1, int64_t b0, int64_t b1) {
    build_dp();
    int W = n + 1;
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = t0; i < t1; ++i) {
        word_t t = or_unrank((uint64_t)i, W);
        uint32_t acc = main_in[i];
        word_t pred[OR_MAXW * 2 + 4];
        int np = 0;
        if (p == 1) np = or_reverse_boundary_preds(t, W, pred, OR_MAXW * 2 + 4);
        else {
            switch (mpair(t, p)) {
            case PAIR(VL, VR): pred[np++] = msetpair(t, p, PAIR(VN, VN)); break;
            case PAIR(VN, VR): pred[np++] = msetpair(t, p, PAIR(VR, VN)); break;
            case PAIR(VN, VL): pred[np++] = msetpair(t, p, PAIR(VL, VN)); break;
            default: break;
            }
        }
        for (int k = 0; k < np; ++k) acc = addm(acc, main_in[or_rank(pred[k], W)], mod);
        if (blk_in && mget(t, p) == VN) acc = addm(acc, blk_in[or_rank(mshrink(t, p), W - 1)], mod);
        main_out[i] = acc;
    }
    if (p == 1) {
        for (int64_t i = b0; i < b1; ++i) blk_out[i] = 0;
        return;
    }
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t i = b0; i < b1; ++i) {
        word_t b = or_unrank((uint64_t)i, W - 1);
        word_t pred[OR_MAXW * 2 + 4];
        int np = or_reverse_block_preds(b, W, p, pred, OR_MAXW * 2 + 4);
        uint32_t acc = 0;
        for (int k = 0; k < np; ++k) acc = addm(acc, main_in[or_rank(pred[k], W)], mod);
        blk_out[i] = acc;
    }
}

/* One cell, pull form over reverse predecessors.  Every main target gathers
 * identity + its included-edge predecessors + its blocked re-expansion; every
 * blocked target gathers its reverse_block_predecessors.  Same result as the
 * forward form (addition mod p is order independent). */
void or_cell_pull(int n, int p, uint32_t mod, const uint32_t* main_in,
                  const uint32_t* blk_in, uint32_t* main_out, uint32_t* blk_out) {
    build_dp();
    int W = n + 1;
    or_cell_pull_range(n, p, mod, main_in, blk_in, main_out, blk_out, 0, (int64_t)DP[W][1], 0,
                       (int64_t)DP[W - 1][1]);
}

/* Apply cells p_hi..p_lo in order to (main, blk) in place (scratch buffers
 * allocated internally).  blk_valid=0 means the blocked input of the first
 * cell is zero. */
int or_apply_cells(int n, uint32_t mod, uint32_t* main, uint32_t* blk, int p_hi, int p_lo,
                   int blk_valid, int use_pull) {
    build_dp();
    int W = n + 1;
    size_t nm = DP[W][1], nb = DP[W - 1][1];
    uint32_t* m2 = (uint32_t*)malloc(nm * sizeof(uint32_t));
    uint32_t* b2 = (uint32_t*)malloc((nb ? nb : 1) * sizeof(uint32_t));
    if (!m2 || !b2) { free(m2); free(b2); return -1; }
    for (int p = p_hi; p >= p_lo; --p) {
        const uint32_t* bin = (p == p_hi && !blk_valid) ? NULL : blk;
        if (use_pull) or_cell_pull(n, p, mod, main, bin, m2, b2);
        else or_cell_forward(n, p, mod, main, bin, m2, b2);
        memcpy(main, m2, nm * sizeof(uint32_t));
        memcpy(blk, b2, nb * sizeof(uint32_t));
    }
    free(m2); free(b2);
    return 0;
}

/* Whole residue: fullmate_dropN_batch.cu:508-514.  Returns the residue or
 * sets *err. */
uint32_t or_solve(int n, uint32_t mod, int use_pull, int* err) {
    build_dp();
    *err = 0;
    if (n < 1 || n + 1 > OR_MAXW || mod < 2) { *err = 1; return 0; }
    int W = n + 1;
    size_t nm = DP[W][1], nb = DP[W - 1][1];
    uint32_t* main = (uint32_t*)calloc(nm, sizeof(uint32_t));
    uint32_t* blk = (uint32_t*)calloc(nb ? nb : 1, sizeof(uint32_t));
    if (!main || !blk) { free(main); free(blk); *err = 2; return 0; }
    main[or_rank((word_t)VR << (2 * (W - 1)), W)] = 1 % mod;
    for (int row = 0; row < W; ++row)
        if (or_apply_cells(n, mod, main, blk, W - 1, 1, 0, use_pull)) { *err = 2; break; }
    uint32_t ans = main[or_rank((word_t)VR, W)];
    free(main); free(blk);
    return ans;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- shard semantics (windowed), for checking one GPU's share ----
 * Synthetic residues of words that live on other shards: a seeded hash of the
 * global index (blocked words in their own domain, bit 62), reduced mod p --
 * the same definition the product's shard sweep uses.  */
uint32_t or_remote_residue(uint64_t seed, uint64_t index, int j, uint32_t mod) {
    uint32_t h = (uint32_t)index * 0x9E3779B1u ^ ((uint32_t)(index >> 32) + (uint32_t)seed) * 0x85EBCA77u ^
                 ((uint32_t)j + (uint32_t)(seed >> 32)) * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    h *= 0x297A2D39u;
    h ^= h >> 15;
    return (uint32_t)(((uint64_t)h * mod) >> 32);
}

/* Cells p_hi..p_lo on a shard: main window [m0, m0+M), blocked words
 * [b0, b0+B).  Window values start as the synthetic residues of their global
 * indices (modulus index j); reads outside the main window return the
 * synthetic residue.  Each cell is the pull form of the reference recurrence
 * over the words of the window (main targets gather identity + included-edge
 * predecessors + blocked re-expansion; blocked targets gather
 * reverse_block_predecessors).  The blocked input of cell W-1 is zero; cell 1
 * leaves the blocked words as they were.  Returns -1 if a blocked read falls
 * outside [b0, b0+B) (a shard must be whole prefix classes). */
int or_shard_cells(int n, uint32_t mod, int j, uint64_t seed, uint64_t b0, uint64_t B, uint64_t m0,
                   uint64_t M, int p_hi, int p_lo, uint32_t* main_win, uint32_t* blk_win) {
    build_dp();
    const int W = n + 1;
    const uint64_t dom = 1ull << 62;
    for (uint64_t i = 0; i < M; ++i) main_win[i] = or_remote_residue(seed, m0 + i, j, mod);
    for (uint64_t b = 0; b < B; ++b) blk_win[b] = or_remote_residue(seed, dom + b0 + b, j, mod);
    uint32_t* nm = (uint32_t*)malloc((M ? M : 1) * sizeof(uint32_t));
    uint32_t* nb = (uint32_t*)malloc((B ? B : 1) * sizeof(uint32_t));
    if (!nm || !nb) { free(nm); free(nb); return -2; }
    int bad = 0;
    for (int p = p_hi; p >= p_lo; --p) {
        const int read_blk = (p != W - 1);
#define OLD(ix) (((uint64_t)(ix) - m0 < M) ? main_win[(uint64_t)(ix) - m0] : or_remote_residue(seed, (ix), j, mod))
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : bad)
        for (int64_t ii = 0; ii < (int64_t)M; ++ii) {
            const uint64_t i = (uint64_t)ii;
            word_t t = or_unrank(m0 + i, W);
            uint32_t acc = main_win[i];
            word_t pred[OR_MAXW * 2 + 4];
            int np = 0;
            if (p == 1) np = or_reverse_boundary_preds(t, W, pred, OR_MAXW * 2 + 4);
            else {
                switch (mpair(t, p)) {
                case PAIR(VL, VR): pred[np++] = msetpair(t, p, PAIR(VN, VN)); break;
                case PAIR(VN, VR): pred[np++] = msetpair(t, p, PAIR(VR, VN)); break;
                case PAIR(VN, VL): pred[np++] = msetpair(t, p, PAIR(VL, VN)); break;
                default: break;
                }
            }
            for (int q = 0; q < np; ++q) acc = addm(acc, OLD(or_rank(pred[q], W)), mod);
            if (read_blk && mget(t, p) == VN) {
                uint64_t bi = or_rank(mshrink(t, p), W - 1);
                if (bi - b0 >= B) bad = 1;
                else acc = addm(acc, blk_win[bi - b0], mod);
            }
            nm[i] = acc;
        }
        if (p >= 2) {
#pragma omp parallel for schedule(dynamic, 1024)
            for (int64_t bb = 0; bb < (int64_t)B; ++bb) {
                const uint64_t b = (uint64_t)bb;
                word_t bw = or_unrank(b0 + b, W - 1);
                word_t pred[OR_MAXW * 2 + 4];
                int np = or_reverse_block_preds(bw, W, p, pred, OR_MAXW * 2 + 4);
                uint32_t acc = 0;
                for (int q = 0; q < np; ++q) acc = addm(acc, OLD(or_rank(pred[q], W)), mod);
                nb[b] = acc;
            }
            memcpy(blk_win, nb, B * sizeof(uint32_t));
        }
#undef OLD
        memcpy(main_win, nm, M * sizeof(uint32_t));
    }
    free(nm);
    free(nb);
    return bad ? -1 : 0;
}
