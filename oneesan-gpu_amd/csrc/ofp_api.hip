// ofp_api.hip -- host runtime behind include/ofp.h.
//
// One context = one GPU, one non-blocking stream (default priority), the
// height tables on the device, and one device-resident frontier state
// (main + blocked words, k moduli interleaved per state).  A solve initialises
// the state (reference: fullmate_dropN_batch.cu:508-511), runs W rows of cells
// p = W-1..1 -- rows whose cells all sweep the whole state are captured once in
// a hipGraph and replayed, the first rows and the last two run height-capped
// cells directly -- and reads main[rank(R)] = main[0]
// (fullmate_dropN_batch.cu:514).  Jobs (serving) hold their own state and
// stream; shard sweeps run one GPU's window of an n=27 row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ofp.h"
#include "ofp_exchange_impl.h"
#include "ofp_internal.h"
#include "ofp_pair_impl.h"

using namespace ofp;

// One device-resident frontier state: main + blocked arrays for (n, moduli),
// the stream its sweeps run on, and the row graph captured for it.
struct Sweep {
    uint32_t* main = nullptr;
    uint32_t* blk = nullptr;
    size_t main_cap = 0, blk_cap = 0;  // bytes allocated
    uint32_t* mods_dev = nullptr;
    int mods_cap = 0;
    int n = 0, W = 0, K = 0, Kp = 0;  // k moduli, interleave stride Kp
    int layout = -1;                  // layout the arrays are written in
    std::vector<uint32_t> mods;
    uint64_t nmain = 0, nblk = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t row_exec = nullptr;
    hipGraph_t row_graph = nullptr;
    std::string row_key;
    int row_launches = 0;      // kernel launches in the captured row graph
    bool async_alloc = false;  // buffers come from the stream-ordered allocator (jobs)
    bool fixed = false;        // buffers are slices of the library's static device array (no graphs)
    int row = -1;              // rows swept since the initial state; -1: state of unknown origin
    std::vector<hipEvent_t> seg_ev;  // timing events of row segments (ofp_row_sweep)
    // pair passes: two more streams for the launches of one pass that may run
    // side by side (created on first use, not for the static sweep)
    hipStream_t par[2] = {};
    hipEvent_t pfork = nullptr, pjoin[2] = {};
};

struct ofp_ctx {
    int device = 0;
    int cus = 256;
    int threads = 256;
    int blocks_per_cu = 32;
    int use_graphs = 1;
    int layout = 1;  // 0: 64-state tiles, 1: state-major with even stride
    int variant = 2;         // kernel variant (cell_kernel V parameter), set by the autotuner
    int low_cells = 0;       // cells p <= low_cells launch at most low_kc moduli at a time
    int low_kc = kMaxKC;
    size_t reserve = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* dp32 = nullptr;  // packed uint32 height table (n <= 22)
    void* dp64 = nullptr;  // packed uint64 height table
    void* dpc32 = nullptr; // packed capped tables, caps 1..kMaxCap (uint32, uint64)
    void* dpc64 = nullptr;
    void* ct32 = nullptr;  // pair-pass class tables, every (p, d) (32-bit ranks)
    void* ct64 = nullptr;  // the same with 64-bit entries (n >= 23, or 64-bit ranks forced)
    int pair_rows = 1;     // 1: full-state rows run as pair passes (two cells per pass, in place)
    int pair_kc = kMaxKC;  // moduli per pair-pass launch (even, <= 16)
    uint32_t pair_mask = 0;  // bit p: cells p and p-1 run as one pair pass; 0: the library's default
    int pair_variant = 0;    // pair-kernel variant (PairLaunch::variant)
    int64_t group_pair_mask = -1;  // group sweeps: bit p pairs cells p, p-1; -1: default_group_pair_mask
    int group_streams = 3;   // streams a group sweep deals each step's group launches to (1..4)
    hipStream_t gstream[3] = {};             // the extra streams (created on first use)
    hipEvent_t gfork = nullptr, gjoin[3] = {};
    int cap_percent = -1;  // height caps: cap a cell when it then sweeps <= this % of the blocked words
                           // (-1: 75 at one modulus, else 85)
    int xcd_shift = 5;     // chunk pieces per XCD: 2^xcd_shift chunks, < 0 = one eighth each
    int sorted_from = -1;  // full-state cells p >= sorted_from run the sorted kernel (0: none,
                           // -1: from cell 6 for launches of 8 or more moduli, else none)
    int index_bits = 0;    // 0: 32-bit ranks for W <= 23, 64-bit above; 64: 64-bit ranks always
    int max_batch = 0;     // moduli per sweep of ofp_solve (0: as many as fit HBM)
    char* win = nullptr;   // window buffer of shard and group sweeps, reused from call to call
    size_t win_bytes = 0;
    // The LOW window the last group sweep left in `win` (what ofp_group_exchange reads).
    struct LowWindow {
        bool valid = false;
        int n = 0, k = 0, low = 0;
        uint32_t first = 0, groups = 0;
        uint32_t* main = nullptr;
        uint32_t* blk = nullptr;
        char* tables = nullptr;
        std::vector<uint64_t> mc, bc;  // words per group
    } last_low;
    char* stage = nullptr;  // exchange buffer: plan tables and staging buckets, reused
    size_t stage_bytes = 0;
    void* scratch = nullptr;  // device address of the static state array on this device
    Sweep fixed;           // ofp_solve_static's state, inside that array
    std::string err;
    Sweep st;              // the context's own state (solve / prepare / row_sweep)
    int jobs = 0;          // live jobs (serving)
};

struct ofp_job {
    ofp_ctx* ctx = nullptr;
    Sweep sw;
    int rows_enqueued = 0;
    hipEvent_t done = nullptr;
};

// The static device array behind ofp_solve_static: one per device the library
// runs on, held until the process ends.  Host code never passes `scratch`
// itself to a kernel or a copy (on the host the name only identifies the
// symbol): the device address comes from hipGetSymbolAddress, once per device
// (static_scratch), and only that address is used.
constexpr size_t kStaticStateBytes = 64ull << 20;
static __device__ __align__(256) uint32_t scratch[kStaticStateBytes / sizeof(uint32_t)];

namespace {

const char* kVersion = "ofp-mi355x 0.2 (per-cell in-place orbit sweep, gfx950)";

ofp_status fail(ofp_ctx* c, ofp_status code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}
ofp_status hip_fail(ofp_ctx* c, hipError_t e, const char* what) {
    return fail(c, OFP_E_RUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}
#define OFP_HIP(c, expr, what)                             \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(c, e_, what); \
    } while (0)

uint64_t dp_at(int w, int h) {
    static uint64_t t[kDpRows][kDpCols];
    static bool ready = false;
    if (!ready) { build_height_table(t); ready = true; }
    return t[w][h];
}

bool counts(int n, uint64_t* m, uint64_t* b) {
    if (n < 1 || n > OFP_MAX_N) return false;
    *m = dp_at(n + 1, 1);
    *b = dp_at(n, 1);
    return true;
}

uint64_t state_bytes(int layout, int n, int k) {
    uint64_t nm, nb;
    if (!counts(n, &nm, &nb)) return 0;
    return (padded_states(nm) + padded_states(nb ? nb : 1)) * (uint64_t)interleave_stride(layout, k) * 4;
}

// Device address of `scratch` on `device`, which must be current; looked up
// once per device and kept.
hipError_t static_scratch(int device, void** out) {
    static std::mutex mu;
    static void* addr[64] = {};
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *out = addr[device];
    return hipSuccess;
}

void drop_graph(Sweep& s) {
    if (s.row_exec) (void)hipGraphExecDestroy(s.row_exec);
    if (s.row_graph) (void)hipGraphDestroy(s.row_graph);
    s.row_exec = nullptr;
    s.row_graph = nullptr;
    s.row_key.clear();
}

// Device buffers of a sweep: jobs use the stream-ordered allocator so that
// creating or retiring one job never waits for the kernels of the others.
hipError_t sweep_alloc(Sweep& s, void** p, size_t bytes) {
    if (s.async_alloc) {
        hipError_t e = hipMallocAsync(p, bytes, s.stream);
        if (e != hipErrorNotSupported) return e;
        (void)hipGetLastError();
        s.async_alloc = false;
    }
    return hipMalloc(p, bytes);
}
void sweep_free(Sweep& s, void* p) {
    if (!p) return;
    if (s.async_alloc) (void)hipFreeAsync(p, s.stream);
    else (void)hipFree(p);
}

void free_sweep(Sweep& s) {
    drop_graph(s);
    for (hipEvent_t e : s.seg_ev) (void)hipEventDestroy(e);
    s.seg_ev.clear();
    for (int q = 0; q < 2; ++q) {
        if (s.par[q]) (void)hipStreamDestroy(s.par[q]);
        if (s.pjoin[q]) (void)hipEventDestroy(s.pjoin[q]);
        s.par[q] = nullptr;
        s.pjoin[q] = nullptr;
    }
    if (s.pfork) (void)hipEventDestroy(s.pfork);
    s.pfork = nullptr;
    sweep_free(s, s.main);
    sweep_free(s, s.blk);
    sweep_free(s, s.mods_dev);
    s.main = s.blk = s.mods_dev = nullptr;
    s.main_cap = s.blk_cap = 0;
    s.mods_cap = 0;
    s.layout = -1;
}

int grid_blocks(const ofp_ctx* c, uint64_t nblk) {
    uint64_t chunks = (nblk + c->threads - 1) / c->threads;
    uint64_t cap = (uint64_t)c->cus * c->blocks_per_cu;
    uint64_t b = std::min<uint64_t>(chunks, cap);
    b = (b + 7) / 8 * 8;
    return (int)std::max<uint64_t>(b, 8);
}

// The window buffer of shard and group sweeps: kept from call to call, so that
// streaming windows of one size through HBM neither frees a touched buffer nor
// touches fresh pages per window (each costs about a second per 100 GB).
// Grown on demand; a full-state sweep gives it up (release_window) before it
// sizes its own state.
ofp_status window_buffer(ofp_ctx* c, size_t bytes, char** out) {
    bytes = (bytes + 255) / 256 * 256;
    if (c->win && c->win_bytes >= bytes) {
        *out = c->win;
        return OFP_OK;
    }
    if (c->win) {
        OFP_HIP(c, hipFree(c->win), "free window buffer");
        c->win = nullptr;
        c->win_bytes = 0;
    }
    size_t fr = 0, tot = 0;
    OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    if (bytes + c->reserve > fr && c->stage) {
        // the staging of the last exchange gives way to a larger window
        OFP_HIP(c, hipFree(c->stage), "free exchange buffer");
        c->stage = nullptr;
        c->stage_bytes = 0;
        OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    }
    if (bytes + c->reserve > fr) return fail(c, OFP_E_MEMORY, "window does not fit HBM");
    // 1/16 headroom where it fits: windows of one stream differ by a few percent,
    // and every regrowth frees touched pages and touches fresh ones
    const size_t slack = bytes + c->reserve + bytes / 16 <= fr ? bytes / 16 / 256 * 256 : 0;
    OFP_HIP(c, hipMalloc(&c->win, bytes + slack), "window buffer");
    c->win_bytes = bytes + slack;
    *out = c->win;
    return OFP_OK;
}

void release_window(ofp_ctx* c) {
    if (c->win) (void)hipFree(c->win);
    c->win = nullptr;
    c->win_bytes = 0;
    if (c->stage) (void)hipFree(c->stage);
    c->stage = nullptr;
    c->stage_bytes = 0;
    c->last_low.valid = false;
}

// Carve consecutive 256-byte aligned pieces out of a window buffer.
struct Carve {
    char* base;
    size_t used = 0;
    template <typename T>
    T* take(size_t bytes) {
        T* p = reinterpret_cast<T*>(base + used);
        used += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return p;
    }
};

// Ensure device buffers of s for (n, moduli) in the context's current layout.
// Whether n in the context's settings runs its full rows as pair passes: pair
// mode on, state-major layout, at least two cells (32-bit ranks for n <= 22,
// 64-bit above or when forced).
bool pairs_apply(const ofp_ctx* c, int n) { return c->pair_rows && c->layout == 1 && n >= 2; }

ofp_status ensure_state(ofp_ctx* c, Sweep& s, int n, const uint32_t* moduli, int k) {
    uint64_t nm, nb;
    if (!counts(n, &nm, &nb)) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    if (k < 1 || k > 1024) return fail(c, OFP_E_ARG, "number of moduli must be in 1..1024");
    for (int j = 0; j < k; ++j)
        if (moduli[j] < 2) return fail(c, OFP_E_ARG, "modulus must be in [2, 4294967295]");
    const int Kp = interleave_stride(c->layout, k);
    size_t mb = padded_states(nm) * (size_t)Kp * sizeof(uint32_t);
    size_t bb = padded_states(nb ? nb : 1) * (size_t)Kp * sizeof(uint32_t);
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    if (mb > s.main_cap || bb > s.blk_cap) {
        drop_graph(s);
        sweep_free(s, s.main);
        sweep_free(s, s.blk);
        s.main = s.blk = nullptr;
        s.main_cap = s.blk_cap = 0;
        release_window(c);
        size_t fr = 0, tot = 0;
        OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        if (mb + bb + c->reserve > fr)
            return fail(c, OFP_E_MEMORY, "insufficient HBM: need " + std::to_string(mb + bb) +
                                             " bytes, free " + std::to_string(fr));
        OFP_HIP(c, sweep_alloc(s, (void**)&s.main, mb), "allocate main state");
        OFP_HIP(c, sweep_alloc(s, (void**)&s.blk, bb), "allocate blocked state");
        s.main_cap = mb;
        s.blk_cap = bb;
    }
    if (s.mods_cap < k) {
        sweep_free(s, s.mods_dev);
        s.mods_dev = nullptr;
        s.mods_cap = std::max(k, 16);
        OFP_HIP(c, sweep_alloc(s, (void**)&s.mods_dev, s.mods_cap * sizeof(uint32_t)), "allocate moduli");
    }
    s.mods.assign(moduli, moduli + k);
    // stream-ordered copy from pageable memory (staged by the runtime before returning)
    OFP_HIP(c, hipMemcpyAsync(s.mods_dev, s.mods.data(), k * sizeof(uint32_t), hipMemcpyHostToDevice, s.stream),
            "copy moduli");
    OFP_HIP(c, hipStreamSynchronize(s.stream), "copy moduli sync");
    s.n = n;
    s.W = n + 1;
    s.K = k;
    s.Kp = Kp;
    s.layout = c->layout;
    s.nmain = nm;
    s.nblk = nb;
    return OFP_OK;
}

// Height caps.  H(w) is the largest height of a word w (the entry height 1
// included).  A thread of cell p reads and writes only words with H >= H(b) of
// its blocked word b (its orbit words, its closure sources and b itself), so:
//  * forward: if every word with H > c holds zero before the cell, the threads
//    with H(b) > c would only add zeros to zeros;
//  * backward: if no word with H > c after the cell can reach the answer word,
//    the outputs of threads with H(b) > c are never read by anything that does;
// either way they can be skipped.  Bounds, for a sweep from the initial state
// (rows r = 0..W-1, cells p = W-1..1):
//  * forward, before cell p of row r: H <= 1 in row 0; in row r >= 1, H <= r at
//    p >= W - r and H <= r + 1 below (a frontier word nests one more arc level
//    only after a further grid row; the kink enters the new level at cell W-r);
//  * backward, after cell p: in the last row only words with H <= 1 still reach
//    the answer (an all-N word with a single R at position 0); in row W-2,
//    after cell p (p >= 2) only those with H <= 2*ceil((p-1)/2)+1, after cell 1
//    only H <= 1.
// Both bounds were checked against the oracle's cell-by-cell sweep at n = 6..14
// (all zeros forward; backward, randomising every word above the bound leaves
// the answer unchanged), and the solves they speed up are checked end to end.
// Returns 0 when the cell runs uncapped.
int cell_height_cap(int W, int r, int p) {
    if (r < 0 || r >= W || p < 1 || p > W - 1) return 0;
    int cap = (r == 0) ? 1 : (p >= W - r ? r : r + 1);
    if (r == W - 1) cap = 1;
    else if (r == W - 2) cap = std::min(cap, p == 1 ? 1 : 2 * (p / 2) + 1);
    return cap <= kMaxCap ? cap : 0;
}

// Blocked words (width W-1, entered at height 1) whose heights stay <= cap.
uint64_t capped_blocked_count(int W, int cap) {
    static thread_local uint64_t d[kDpRows][kDpCols];
    build_capped_table(cap, d);
    return d[W - 1][1];
}

// Piece size (log2 chunks) of one launch: the context's, lowered until the launch
// has at least 64 pieces.  With fewer pieces than XCDs the pieces land on a few
// XCDs and their workgroups walk the chunks one after another (a launch of 8
// chunks in one 32-chunk piece ran on one workgroup); 64 pieces keep the XCDs
// within one piece of each other.
int piece_shift(int shift, uint64_t nwords, int threads) {
    const uint64_t chunks = (nwords + threads - 1) / threads;
    while (shift > 0 && (chunks >> shift) < 64) --shift;
    return shift;
}

// The cap a cell runs with: the bound above, when the capped kernel has few
// enough words to pay for its longer decode (<= percent% of the blocked words).
int effective_cap(int W, int r, int p, uint64_t nblk, uint64_t* count, int percent = 85) {
    const int cap = cell_height_cap(W, r, p);
    if (!cap || percent <= 0) return 0;
    const uint64_t nc = capped_blocked_count(W, cap);
    if (nc * 100 > nblk * (uint64_t)percent) return 0;
    if (count) *count = nc;
    return cap;
}

// The height-cap threshold of a sweep: the context's, else 75 % at one modulus
// (index-bound: the capped kernel's longer decode weighs more) and 85 % above.
int cap_pct(const ofp_ctx* c, const Sweep& s) {
    return c->cap_percent >= 0 ? c->cap_percent : (s.K == 1 ? 75 : 85);
}

// Enqueue cells p_hi..p_lo of s on s.stream; row >= 0 names the row of a sweep
// from the initial state (height caps apply), -1 an arbitrary state.
ofp_status enqueue_cells(ofp_ctx* c, Sweep& s, int p_hi, int p_lo, bool first_reads_blk, int* launches,
                         int row = -1, hipEvent_t* cell_ev = nullptr, int* cell_kind = nullptr) {
    const bool wide = s.W > 23 || c->index_bits == 64;
    const int blocks = grid_blocks(c, s.nblk);
    const int pct = cap_pct(c, s);
    const bool caps = pct > 0 && row >= 0 && s.layout == 1;
    for (int p = p_hi; p >= p_lo; --p) {
        uint64_t ncap = 0;
        const int cap = caps ? effective_cap(s.W, row, p, s.nblk, &ncap, pct) : 0;
        // per-cell timing (ofp_time_cells): events around each cell's launches
        if (cell_ev) OFP_HIP(c, hipEventRecord(cell_ev[2 * (p_hi - p)], s.stream), "event record");
        // moduli per launch: all (up to 16) at once, or fewer on the low cells,
        // whose scattered closure reads are latency-bound (fewer registers per
        // thread -> more waves in flight).  Chunks keep an even offset.
        int step = kMaxKC;
        if (p <= c->low_cells && c->low_kc < step) step = c->low_kc;
        if (s.layout == 1 && s.K > 1 && (step & 1)) step = std::max(2, step - 1);
        for (int off = 0; off < s.K; off += step) {
            CellLaunch a{};
            a.main = s.main;
            a.blk = s.blk;
            a.dp = wide ? c->dp64 : c->dp32;
            a.nblk = s.nblk;
            a.first = 0;
            a.layout = s.layout;
            // the sorted assignment (variant 4) on the cells from sorted_from up
            const int sf = c->sorted_from >= 0 ? c->sorted_from
                                                : (std::min(step, s.K - off) >= 8 ? 6 : 0);
            a.variant = (sf > 0 && p >= sf && s.layout == 1) ? 4 : c->variant;
            a.xcd_shift = piece_shift(c->xcd_shift, cap ? ncap : s.nblk, c->threads);
            a.W = s.W;
            a.p = p;
            a.read_blk = (p == p_hi) ? (first_reads_blk ? 1 : 0) : 1;
            a.K = s.Kp;
            a.off = off;
            a.kc = std::min(step, s.K - off);
            for (int j = 0; j < a.kc; ++j) a.mod[j] = s.mods[off + j];
            hipError_t e;
            if (cap) {
                if (ncap == 0) continue;
                a.nblk = ncap;
                const size_t one = (wide ? 2 * sizeof(uint64_t) : 2 * sizeof(uint32_t)) * kDpRows * kDpCols;
                a.dpc = (const char*)(wide ? c->dpc64 : c->dpc32) + one * (cap - 1);
                e = launch_cell_capped(a, wide, c->threads, grid_blocks(c, ncap), s.stream);
            } else {
                e = launch_cell(a, wide, c->threads, blocks, s.stream);
            }
            if (e != hipSuccess) return hip_fail(c, e, "cell kernel launch");
            if (launches) ++*launches;
            if (cell_kind) cell_kind[p_hi - p] = a.variant == 4 ? 1 : 0;
        }
        if (cell_ev) OFP_HIP(c, hipEventRecord(cell_ev[2 * (p_hi - p) + 1], s.stream), "event record");
    }
    return OFP_OK;
}

// True when s's full rows run as pair passes.
bool pair_ready(const ofp_ctx* c, const Sweep& s) { return pairs_apply(c, s.n) && s.layout == 1; }

// Cells p_hi..p_lo of s, walking down from p_hi: where bit p of the context's
// pair mask is set and p-1 >= p_lo, cells p and p-1 run as one pair pass (in
// place, one launch per window height change d, the classes that read other
// classes -- d = -1, 0, 1 -- first), otherwise cell p runs alone with the
// per-cell kernel.  The first cell reads the blocked input iff first_reads_blk.
// With cell_ev, events bracket each step: a pair's first cell slot (slot
// p_hi - p, kind 2) gets its time and the second an empty interval (kind 3).
// The library's default pairing: every pair except the one over cells 4 and 3,
// which runs as two launches (n=22: 125 against 60 + 66 ms at 14 moduli, 27.4
// against 12.8 + 13.3 ms at one modulus; the other passes are 1.2-1.5x faster
// than their two cells).
uint32_t default_pair_mask(int) { return ~((1u << 4) | (1u << 3)); }
// Group sweeps (4 moduli, n=27 windows over three streams): pair passes over
// every pair of cells from 7 up; cells 6..1 one per launch, where the pair
// pass over a group is no faster than its two cells (6, 5: 781 against 780 ms
// per share row) or slower (2, 1: 1127 against 957 ms).
uint32_t default_group_pair_mask() { return ~0x7fu; }

// The pair passes' extra streams and events of a sweep (created once, before
// any row is captured into a graph).
ofp_status pass_streams(ofp_ctx* c, Sweep& s) {
    if (s.pfork) return OFP_OK;  // set last: all of them exist
    for (int q = 0; q < 2; ++q) {
        if (!s.par[q]) OFP_HIP(c, hipStreamCreateWithFlags(&s.par[q], hipStreamNonBlocking), "pair pass stream");
        if (!s.pjoin[q]) OFP_HIP(c, hipEventCreateWithFlags(&s.pjoin[q], hipEventDisableTiming), "pair pass event");
    }
    OFP_HIP(c, hipEventCreateWithFlags(&s.pfork, hipEventDisableTiming), "pair pass event");
    return OFP_OK;
}

ofp_status enqueue_pairs(ofp_ctx* c, Sweep& s, int p_hi, int p_lo, bool first_reads_blk, int* launches,
                         hipEvent_t* cell_ev = nullptr, int* cell_kind = nullptr) {
    const int W = s.W;
    const bool wide = s.W > 23 || c->index_bits == 64;
    const size_t tb = class_table_bytes(wide);
    const uint32_t mask = c->pair_mask ? c->pair_mask : default_pair_mask(std::min(s.K, c->pair_kc));
    for (int p = p_hi; p >= p_lo;) {
        const int slot = p_hi - p;
        const bool reads_blk = p == p_hi ? first_reads_blk : true;
        if (!(p - 1 >= p_lo && p >= 2 && ((mask >> p) & 1u))) {
            ofp_status st = enqueue_cells(c, s, p, p, reads_blk, launches, -1, cell_ev ? cell_ev + 2 * slot : nullptr,
                                          cell_kind ? cell_kind + slot : nullptr);
            if (st) return st;
            --p;
            continue;
        }
        if (cell_ev) OFP_HIP(c, hipEventRecord(cell_ev[2 * slot], s.stream), "event record");
        for (int off = 0; off < s.Kp; off += c->pair_kc) {
            const int kc = std::min(c->pair_kc, s.Kp - off);
            static const int kOrder[7] = {-1, 0, 1, -3, -2, 2, 3};
            // The readers (d = -1, 0, 1) never read each other's changed words and
            // the d = -2, 2 classes read only their own, so each of the two groups
            // of launches runs side by side on three streams (fork before, join
            // after): their ramp-ups and tails overlap.
            const bool par = !s.fixed;
            if (par)
                if (ofp_status f = pass_streams(c, s)) return f;
            auto fork = [&]() -> ofp_status {
                if (!par) return OFP_OK;
                OFP_HIP(c, hipEventRecord(s.pfork, s.stream), "pair pass fork");
                for (int q = 0; q < 2; ++q) OFP_HIP(c, hipStreamWaitEvent(s.par[q], s.pfork, 0), "pair pass fork");
                return OFP_OK;
            };
            auto join = [&]() -> ofp_status {
                if (!par) return OFP_OK;
                for (int q = 0; q < 2; ++q) {
                    OFP_HIP(c, hipEventRecord(s.pjoin[q], s.par[q]), "pair pass join");
                    OFP_HIP(c, hipStreamWaitEvent(s.stream, s.pjoin[q], 0), "pair pass join");
                }
                return OFP_OK;
            };
            for (int di = 0; di < 7; ++di) {
                const int d = kOrder[di];
                if (di == 3)
                    if (ofp_status f = join()) return f;
                if (di == 0 || di == 4)
                    if (ofp_status f = fork()) return f;
                // d = -3 / +3: the class is one word, RRR or LLL, which neither cell
                // changes (no target at (p, p-1) nor at (p-1, p-2)): nothing to do
                if (d == -3 || d == 3) continue;
                const uint64_t cnt = pair_class_count(W, p, d);
                if (!cnt) continue;
                pair::PairLaunch a{};
                a.main_in = s.main;
                a.blk_in = s.blk;
                a.main_out = s.main;
                a.blk_out = s.blk;
                a.dp = wide ? c->dp64 : c->dp32;
                a.ct = static_cast<const char*>(wide ? c->ct64 : c->ct32) + tb * (size_t)(p * 7 + d + 3);
                a.classes = cnt;
                a.W = W;
                a.p = p;
                a.d = d;
                a.read_blk = reads_blk ? 1 : 0;
                a.K = s.Kp;
                a.off = off;
                a.kc = kc;
                // lanes past the k moduli (stride padding) compute with a valid modulus
                for (int jj = 0; jj < kc; ++jj) a.mod[jj] = s.mods[off + jj < s.K ? off + jj : 0];
                a.xcd_shift = piece_shift(c->xcd_shift, cnt, c->threads);
                a.variant = c->pair_variant;
                hipStream_t ls = !par ? s.stream : (d == -1 || d == -2) ? s.par[0] : (d == 1 ? s.par[1] : s.stream);
                hipError_t e = launch_pair(a, wide, c->threads, grid_blocks(c, cnt), ls);
                if (e != hipSuccess) return hip_fail(c, e, "pair pass launch");
                if (launches) ++*launches;
            }
            if (ofp_status f = join()) return f;
        }
        if (cell_ev) {
            OFP_HIP(c, hipEventRecord(cell_ev[2 * slot + 1], s.stream), "event record");
            OFP_HIP(c, hipEventRecord(cell_ev[2 * (slot + 1)], s.stream), "event record");
            OFP_HIP(c, hipEventRecord(cell_ev[2 * (slot + 1) + 1], s.stream), "event record");
        }
        if (cell_kind) {
            cell_kind[slot] = 2;      // the pass's time
            cell_kind[slot + 1] = 3;  // its second cell: no time of its own
        }
        p -= 2;
    }
    return OFP_OK;
}

// `rows` uncapped rows (cells W-1..1; the blocked input of cell W-1 is zero),
// via a hipGraph captured once per (state, launch shape) and replayed.
ofp_status full_rows(ofp_ctx* c, Sweep& s, int rows, int* launches) {
    const bool pairs = pair_ready(c, s);
    // one whole-state row: pair passes where they apply, else one launch per cell
    auto row = [&](int* l) {
        return pairs ? enqueue_pairs(c, s, s.W - 1, 1, false, l) : enqueue_cells(c, s, s.W - 1, 1, false, l);
    };
    if (!c->use_graphs || s.fixed) {
        for (int r = 0; r < rows; ++r) {
            ofp_status st = row(launches);
            if (st) return st;
        }
        return OFP_OK;
    }
    char key[320];
    snprintf(key, sizeof key, "%p/%p/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%d/%u", (void*)s.main, (void*)s.blk, s.n, s.K,
             c->threads, c->blocks_per_cu, s.layout, c->low_cells, c->low_kc, c->variant * 100 + c->sorted_from,
             c->xcd_shift, c->index_bits, pairs ? 1 + c->pair_variant : 0, c->pair_kc, c->pair_mask);
    std::string k = key;
    for (uint32_t m : s.mods) k += "/" + std::to_string(m);
    if (k != s.row_key || !s.row_exec) {
        drop_graph(s);
        if (pairs)
            if (ofp_status f = pass_streams(c, s)) return f;
        OFP_HIP(c, hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "begin capture");
        int captured = 0;
        ofp_status st = row(&captured);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s.stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return hip_fail(c, e, "end capture");
        s.row_graph = g;
        OFP_HIP(c, hipGraphInstantiate(&s.row_exec, g, nullptr, nullptr, 0), "graph instantiate");
        s.row_key = k;
        s.row_launches = captured;
    }
    for (int r = 0; r < rows; ++r) {
        OFP_HIP(c, hipGraphLaunch(s.row_exec, s.stream), "graph launch");
        if (launches) *launches += s.row_launches;
    }
    return OFP_OK;
}

// True when some cell of row r (of a sweep from the initial state) runs capped.
bool row_capped(const ofp_ctx* c, const Sweep& s, int r) {
    const int pct = cap_pct(c, s);
    if (pct <= 0 || r < 0 || s.layout != 1) return false;
    for (int p = s.W - 1; p >= 1; --p)
        if (effective_cap(s.W, r, p, s.nblk, nullptr, pct)) return true;
    return false;
}

// A row with capped cells (row r of a sweep from the initial state): each capped
// cell alone, each run of uncapped cells as full-state cells, paired as in a
// full row where pair passes apply.  (A pair pass is two consecutive full-state
// cells, so the row computes what one launch per cell computes.)
ofp_status enqueue_capped_row(ofp_ctx* c, Sweep& s, int r, int* launches) {
    if (!pair_ready(c, s)) return enqueue_cells(c, s, s.W - 1, 1, false, launches, r);
    const int pct = cap_pct(c, s);
    for (int p = s.W - 1; p >= 1;) {
        const bool reads_blk = p != s.W - 1;
        if (effective_cap(s.W, r, p, s.nblk, nullptr, pct)) {
            ofp_status st = enqueue_cells(c, s, p, p, reads_blk, launches, r);
            if (st) return st;
            --p;
            continue;
        }
        int q = p;
        while (q - 1 >= 1 && !effective_cap(s.W, r, q - 1, s.nblk, nullptr, pct)) --q;
        ofp_status st = enqueue_pairs(c, s, p, q, reads_blk, launches);
        if (st) return st;
        p = q - 1;
    }
    return OFP_OK;
}

// Kernel time split between rows that run only uncapped cells (the full-state
// kernel, replayed from the row graph) and rows with capped cells.
struct RowTimes {
    double full_s = 0, capped_s = 0;
    int full_launches = 0, capped_launches = 0;
};

// `rows` rows from the sweep's current row: rows with capped cells are launched
// directly, runs of uncapped rows from the row graph.  With `times`, each run
// is bracketed by events and the call waits for the last one.
ofp_status run_rows(ofp_ctx* c, Sweep& s, int rows, int* launches, RowTimes* times = nullptr) {
    struct Seg {
        bool capped;
        int launches;
    };
    std::vector<Seg> segs;
    auto event = [&](size_t i) -> hipEvent_t {
        while (s.seg_ev.size() <= i) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            s.seg_ev.push_back(e);
        }
        return s.seg_ev[i];
    };
    int done = 0;
    while (done < rows) {
        const int r = s.row >= 0 ? s.row + done : -1;
        const bool capped = row_capped(c, s, r);
        int n = 1;
        if (!capped)
            while (done + n < rows && !row_capped(c, s, r < 0 ? -1 : r + n)) ++n;
        if (times) {
            hipEvent_t e = event(2 * segs.size());
            if (!e) return fail(c, OFP_E_RUNTIME, "event create");
            OFP_HIP(c, hipEventRecord(e, s.stream), "event record");
        }
        int l = 0;
        ofp_status st = capped ? enqueue_capped_row(c, s, r, &l) : full_rows(c, s, n, &l);
        if (st) return st;
        if (times) {
            hipEvent_t e = event(2 * segs.size() + 1);
            if (!e) return fail(c, OFP_E_RUNTIME, "event create");
            OFP_HIP(c, hipEventRecord(e, s.stream), "event record");
        }
        segs.push_back({capped, l});
        if (launches) *launches += l;
        done += n;
    }
    if (s.row >= 0) s.row += rows;
    if (times && !segs.empty()) {
        OFP_HIP(c, hipEventSynchronize(s.seg_ev[2 * segs.size() - 1]), "row sweep sync");
        for (size_t i = 0; i < segs.size(); ++i) {
            float ms = 0;
            OFP_HIP(c, hipEventElapsedTime(&ms, s.seg_ev[2 * i], s.seg_ev[2 * i + 1]), "event time");
            if (segs[i].capped) {
                times->capped_s += ms * 1e-3;
                times->capped_launches += segs[i].launches;
            } else {
                times->full_s += ms * 1e-3;
                times->full_launches += segs[i].launches;
            }
        }
    }
    return OFP_OK;
}

// Zero main and blocked and place the start word R at position W-1 (rank
// H_DP[W-1][1]) with residue 1 (fullmate_dropN_batch.cu:508-511).
ofp_status init_state(ofp_ctx* c, Sweep& s) {
    OFP_HIP(c, hipMemsetAsync(s.main, 0, padded_states(s.nmain) * (size_t)s.Kp * 4, s.stream), "zero main state");
    // capped cells leave the blocked words above their cap untouched: those must read as zero
    OFP_HIP(c, hipMemsetAsync(s.blk, 0, padded_states(s.nblk ? s.nblk : 1) * (size_t)s.Kp * 4, s.stream),
            "zero blocked state");
    s.row = 0;
    uint64_t start = dp_at(s.W - 1, 1);
    hipError_t e = s.mods_dev ? launch_set_one(s.main, start, s.Kp, s.K, s.layout, s.mods_dev, s.stream)
                              : launch_set_one_values(s.main, start, s.Kp, s.K, s.layout, s.mods.data(), s.stream);
    if (e != hipSuccess) return hip_fail(c, e, "init kernel");
    return OFP_OK;
}

// Split k moduli into sweeps of at most g each (g >= 1): as few sweeps as
// possible, sizes as equal as possible, and even sizes where that costs no
// sweep (the state-major stride is padded to even, so an odd sweep carries
// one idle slot).  Returns the number of sweeps; sizes[] gets the first cap.
int moduli_groups(int k, int g, int* sizes, int cap) {
    if (k < 1 || g < 1) return 0;
    const int groups = (k + g - 1) / g;
    int per = (k + groups - 1) / groups;
    if ((per & 1) && per + 1 <= g && (k + per) / (per + 1) <= groups) ++per;
    int left = k, i = 0;
    while (left > 0) {
        const int now = std::min(per, left);
        if (sizes && i < cap) sizes[i] = now;
        left -= now;
        ++i;
    }
    return i;
}

// Host rank / unrank over the height table (lexicographic, N<R<L, entered at height 1).
uint64_t rank_word(uint64_t w, int width) {
    uint64_t r = 0;
    int h = 1;
    for (int pos = width - 1; pos >= 0; --pos) {
        int v = (int)((w >> (2 * pos)) & 3u);
        if (v >= 1) r += dp_at(pos, h);
        if (v == 2 && h > 0) r += dp_at(pos, h - 1);
        h += (v == 2) - (v == 1);
    }
    return r;
}
uint64_t unrank_word(uint64_t r, int width, int stop_pos, int* h_at_stop) {
    uint64_t w = 0;
    int h = 1;
    for (int pos = width - 1; pos >= 0; --pos) {
        if (pos == stop_pos - 1 && h_at_stop) *h_at_stop = h;
        uint64_t z = dp_at(pos, h);
        if (r < z) continue;
        r -= z;
        uint64_t z2 = h > 0 ? dp_at(pos, h - 1) : 0;
        if (r < z2) { w |= 1ull << (2 * pos); --h; continue; }
        r -= z2;
        w |= 2ull << (2 * pos);
        ++h;
    }
    if (stop_pos == 0 && h_at_stop) *h_at_stop = h;
    return w;
}
// Smallest legal suffix over `len` positions entered at height h (N's, then h R's).
uint64_t min_suffix(int len, int h) {
    uint64_t w = 0;
    for (int pos = 0; pos < h && pos < len; ++pos) w |= 1ull << (2 * pos);
    return w;
}
// First blocked / main index of the prefix class (blocked positions >= lc) that
// contains blocked word r.
void class_start(int W, int lc, uint64_t r, uint64_t* b_first, uint64_t* m_first) {
    int h = 1;
    uint64_t w = unrank_word(r, W - 1, lc, &h);
    uint64_t prefix = w & ~((1ull << (2 * lc)) - 1ull);
    *b_first = rank_word(prefix | min_suffix(lc, h), W - 1);
    *m_first = rank_word((prefix << 2) | min_suffix(lc + 1, h), W);
}

ofp_status read_main(ofp_ctx* c, Sweep& s, uint64_t index, uint32_t* out_k, int capacity) {
    if (!s.main) return fail(c, OFP_E_ARG, "no prepared state");
    if (index >= s.nmain) return fail(c, OFP_E_ARG, "index out of range");
    if (capacity < s.K)
        return fail(c, OFP_E_ARG, "output holds " + std::to_string(capacity) + " residues, the state has " +
                                      std::to_string(s.K) + " moduli");
    OFP_HIP(c, hipStreamSynchronize(s.stream), "sync");
    for (int j = 0; j < s.K; ++j)
        OFP_HIP(c, hipMemcpy(out_k + j, s.main + state_elem(s.layout, index, s.Kp, j), 4, hipMemcpyDeviceToHost),
                "read residue");
    return OFP_OK;
}

}  // namespace

extern "C" {

const char* ofp_version(void) { return kVersion; }

const char* ofp_last_error(const ofp_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

ofp_status ofp_state_counts(int n, uint64_t* main_states, uint64_t* blocked_states) {
    uint64_t m, b;
    if (!main_states || !blocked_states) return OFP_E_ARG;
    if (!counts(n, &m, &b)) return OFP_E_RANGE;
    *main_states = m;
    *blocked_states = b;
    return OFP_OK;
}

ofp_status ofp_create(const ofp_config* cfg, ofp_ctx** out) {
    if (!out) return OFP_E_ARG;
    *out = nullptr;
    ofp_ctx* c = new ofp_ctx();
    if (cfg) {
        c->device = cfg->device;
        if (cfg->threads) c->threads = cfg->threads;
        if (cfg->blocks_per_cu) c->blocks_per_cu = cfg->blocks_per_cu;
        c->reserve = cfg->reserve_bytes;
        c->use_graphs = cfg->use_graphs;
        if (cfg->layout == 1 || cfg->layout == 2) c->layout = cfg->layout - 1;
    }
    if (c->threads < 64 || c->threads > 256 || (c->threads & 63) || c->blocks_per_cu <= 0) {
        delete c;
        return OFP_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || c->device < 0 || c->device >= ndev) {
        delete c;
        return OFP_E_DEVICE;
    }
    auto bad = [&](hipError_t e) {
        if (e == hipSuccess) return false;
        if (c->dp32) (void)hipFree(c->dp32);
        if (c->dp64) (void)hipFree(c->dp64);
        if (c->dpc32) (void)hipFree(c->dpc32);
        if (c->dpc64) (void)hipFree(c->dpc64);
        if (c->ct32) (void)hipFree(c->ct32);
        if (c->ct64) (void)hipFree(c->ct64);
        if (c->ev0) (void)hipEventDestroy(c->ev0);
        if (c->ev1) (void)hipEventDestroy(c->ev1);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return true;
    };
    if (bad(hipSetDevice(c->device))) return OFP_E_DEVICE;
    hipDeviceProp_t prop;
    if (bad(hipGetDeviceProperties(&prop, c->device))) return OFP_E_DEVICE;
    c->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (!c->reserve) {
        // reference default: clamp(total/32, 256 MiB, 8 GiB) (fullmate_dropN_batch.cu:499)
        size_t tot = prop.totalGlobalMem;
        c->reserve = std::min<size_t>(8ull << 30, std::max<size_t>(256ull << 20, tot / 32));
    }
    if (bad(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking))) return OFP_E_RUNTIME;
    c->st.stream = c->stream;
    c->fixed.stream = c->stream;
    c->fixed.fixed = true;
    if (bad(static_scratch(c->device, &c->scratch))) return OFP_E_RUNTIME;
    if (bad(hipEventCreate(&c->ev0)) || bad(hipEventCreate(&c->ev1))) return OFP_E_RUNTIME;
    std::vector<unsigned char> q32(64 * 1024), q64(64 * 1024);
    size_t b32 = pack_height_table(false, q32.data(), q32.size());
    size_t b64 = pack_height_table(true, q64.data(), q64.size());
    if (!b32 || !b64) {
        (void)bad(hipErrorInvalidValue);
        return OFP_E_RUNTIME;
    }
    if (bad(hipMalloc(&c->dp32, b32)) || bad(hipMalloc(&c->dp64, b64))) return OFP_E_RUNTIME;
    if (bad(hipMemcpy(c->dp32, q32.data(), b32, hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(c->dp64, q64.data(), b64, hipMemcpyHostToDevice)))
        return OFP_E_RUNTIME;
    std::vector<unsigned char> c32(512 * 1024), c64(512 * 1024);
    size_t bc32 = pack_capped_tables(false, c32.data(), c32.size());
    size_t bc64 = pack_capped_tables(true, c64.data(), c64.size());
    if (!bc32 || !bc64) {
        (void)bad(hipErrorInvalidValue);
        return OFP_E_RUNTIME;
    }
    if (bad(hipMalloc(&c->dpc32, bc32)) || bad(hipMalloc(&c->dpc64, bc64))) return OFP_E_RUNTIME;
    if (bad(hipMemcpy(c->dpc32, c32.data(), bc32, hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(c->dpc64, c64.data(), bc64, hipMemcpyHostToDevice)))
        return OFP_E_RUNTIME;
    std::vector<unsigned char> t32(class_table_bytes(false) * kMaxW * 7);
    const size_t bt32 = pack_class_tables(false, t32.data(), t32.size());
    if (!bt32) {
        (void)bad(hipErrorInvalidValue);
        return OFP_E_RUNTIME;
    }
    if (bad(hipMalloc(&c->ct32, bt32)) || bad(hipMemcpy(c->ct32, t32.data(), bt32, hipMemcpyHostToDevice)))
        return OFP_E_RUNTIME;
    std::vector<unsigned char> t64(class_table_bytes(true) * kMaxW * 7);
    const size_t bt64 = pack_class_tables(true, t64.data(), t64.size());
    if (!bt64) {
        (void)bad(hipErrorInvalidValue);
        return OFP_E_RUNTIME;
    }
    if (bad(hipMalloc(&c->ct64, bt64)) || bad(hipMemcpy(c->ct64, t64.data(), bt64, hipMemcpyHostToDevice)))
        return OFP_E_RUNTIME;
    *out = c;
    return OFP_OK;
}

ofp_status ofp_set_launch(ofp_ctx* c, int threads, int blocks_per_cu) {
    if (!c) return OFP_E_ARG;
    if (threads == 0) threads = 256;
    if (blocks_per_cu == 0) blocks_per_cu = 32;
    if (threads < 64 || threads > 256 || (threads & 63) || blocks_per_cu < 1 || blocks_per_cu > 64)
        return fail(c, OFP_E_ARG, "threads must be 64..256 in steps of 64, blocks_per_cu 1..64");
    if (threads != c->threads || blocks_per_cu != c->blocks_per_cu) drop_graph(c->st);
    c->threads = threads;
    c->blocks_per_cu = blocks_per_cu;
    return OFP_OK;
}

ofp_status ofp_set_variant(ofp_ctx* c, int variant) {
    if (!c) return OFP_E_ARG;
    if (variant < 0 || variant > 4) return fail(c, OFP_E_ARG, "variant must be 0..4");
    if (variant != c->variant) drop_graph(c->st);
    c->variant = variant;
    return OFP_OK;
}

ofp_status ofp_set_xcd_pieces(ofp_ctx* c, int log2_chunks) {
    if (!c) return OFP_E_ARG;
    if (log2_chunks < -1 || log2_chunks > 24) return fail(c, OFP_E_ARG, "log2_chunks must be -1..24");
    if (log2_chunks != c->xcd_shift) drop_graph(c->st);
    c->xcd_shift = log2_chunks;
    return OFP_OK;
}

ofp_status ofp_set_sorted_cells(ofp_ctx* c, int from_cell) {
    if (!c) return OFP_E_ARG;
    if (from_cell < -1 || from_cell > kMaxW) return fail(c, OFP_E_ARG, "from_cell must be -1..28");
    if (from_cell != c->sorted_from) drop_graph(c->st);
    c->sorted_from = from_cell;
    return OFP_OK;
}

ofp_status ofp_set_height_caps(ofp_ctx* c, int max_percent) {
    if (!c) return OFP_E_ARG;
    if (max_percent < -1 || max_percent > 100) return fail(c, OFP_E_ARG, "max_percent must be -1..100");
    c->cap_percent = max_percent;
    return OFP_OK;
}

int ofp_height_cap(int n, int row, int p, int max_percent, uint64_t* capped_blocked) {
    if (n < 1 || n > OFP_MAX_N) return -1;
    const int W = n + 1;
    uint64_t nm, nb;
    if (!counts(n, &nm, &nb)) return -1;
    uint64_t nc = nb;
    const int cap = effective_cap(W, row, p, nb, &nc, max_percent);
    if (capped_blocked) *capped_blocked = cap ? nc : nb;
    return cap;
}

ofp_status ofp_set_low_cells(ofp_ctx* c, int low_cells, int low_kc) {
    if (!c) return OFP_E_ARG;
    if (low_cells < 0 || low_cells > kMaxW || low_kc < 1 || low_kc > kMaxKC)
        return fail(c, OFP_E_ARG, "low_cells must be 0..28 and low_kc 1..16");
    if (low_cells != c->low_cells || low_kc != c->low_kc) drop_graph(c->st);
    c->low_cells = low_cells;
    c->low_kc = low_kc;
    return OFP_OK;
}

ofp_status ofp_set_layout(ofp_ctx* c, int layout) {
    if (!c) return OFP_E_ARG;
    if (layout != 0 && layout != 1) return fail(c, OFP_E_ARG, "layout must be 0 (64-state tiles) or 1 (state-major)");
    if (layout != c->layout) {
        drop_graph(c->st);
        c->layout = layout;
        c->st.layout = -1;  // current contents no longer readable; prepare again
    }
    return OFP_OK;
}

ofp_status ofp_plan(ofp_ctx* c, int n, int k, ofp_plan_info* out) {
    if (!c || !out) return OFP_E_ARG;
    uint64_t nm, nb;
    if (!counts(n, &nm, &nb)) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    if (k < 1) return fail(c, OFP_E_ARG, "k must be >= 1");
    out->main_states = nm;
    out->blocked_states = nb;
    out->state_bytes = state_bytes(c->layout, n, k);
    const uint64_t W = n + 1;
    out->transitions = W * (W - 1) * (nm + nb) * (uint64_t)k;
    size_t fr = 0, tot = 0;
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    // memory already held by this context's own state can be reused
    out->free_bytes = fr + c->st.main_cap + c->st.blk_cap;
    out->fits = out->state_bytes + c->reserve <= out->free_bytes;
    return OFP_OK;
}

ofp_status ofp_prepare(ofp_ctx* c, int n, const uint32_t* moduli, int k) {
    if (!c || !moduli) return OFP_E_ARG;
    ofp_status s = ensure_state(c, c->st, n, moduli, k);
    if (s) return s;
    s = init_state(c, c->st);
    if (s) return s;
    OFP_HIP(c, hipStreamSynchronize(c->stream), "prepare sync");
    return OFP_OK;
}

ofp_status ofp_row_sweep(ofp_ctx* c, int rows, ofp_stats* st) {
    if (!c || rows < 0) return OFP_E_ARG;
    Sweep& s = c->st;
    if (!s.main || s.layout != c->layout) return fail(c, OFP_E_ARG, "no prepared state");
    int launches = 0;
    auto t0 = std::chrono::steady_clock::now();
    OFP_HIP(c, hipEventRecord(c->ev0, s.stream), "event record");
    RowTimes times;
    ofp_status r = run_rows(c, s, rows, &launches, &times);
    if (r) return r;
    OFP_HIP(c, hipEventRecord(c->ev1, s.stream), "event record");
    OFP_HIP(c, hipEventSynchronize(c->ev1), "row sweep sync");
    float ms = 0;
    OFP_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1), "event time");
    if (st) {
        st->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        st->kernel_s = ms * 1e-3;
        st->full_kernel_s = times.full_s;
        st->full_launches = times.full_launches;
        st->capped_kernel_s = times.capped_s;
        st->capped_launches = times.capped_launches;
        st->main_states = s.nmain;
        st->blocked_states = s.nblk;
        st->rows = rows;
        st->cells_per_row = s.W - 1;
        st->launches = launches;
        st->transitions = (uint64_t)rows * (s.W - 1) * (s.nmain + s.nblk) * (uint64_t)s.K;
    }
    return OFP_OK;
}

ofp_status ofp_read_main(ofp_ctx* c, uint64_t index, uint32_t* out, int capacity) {
    if (!c || !out) return OFP_E_ARG;
    if (c->st.layout != c->layout) return fail(c, OFP_E_ARG, "no prepared state");
    return read_main(c, c->st, index, out, capacity);
}

ofp_status ofp_time_cells(ofp_ctx* c, float* ms_out, int* kind_out, int capacity) {
    if (!c || !ms_out || !kind_out) return OFP_E_ARG;
    Sweep& s = c->st;
    if (!s.main || s.layout != c->layout) return fail(c, OFP_E_ARG, "no prepared state");
    const int cells = s.W - 1;
    if (capacity < cells) return fail(c, OFP_E_ARG, "capacity below the W-1 cells of a row");
    while ((int)s.seg_ev.size() < 2 * cells) {
        hipEvent_t e = nullptr;
        OFP_HIP(c, hipEventCreate(&e), "event create");
        s.seg_ev.push_back(e);
    }
    for (int i = 0; i < cells; ++i) kind_out[i] = 0;
    // row -1: an arbitrary state, every cell over the whole state
    ofp_status st = pair_ready(c, s) ? enqueue_pairs(c, s, s.W - 1, 1, false, nullptr, s.seg_ev.data(), kind_out)
                                     : enqueue_cells(c, s, s.W - 1, 1, false, nullptr, -1, s.seg_ev.data(), kind_out);
    if (st) return st;
    OFP_HIP(c, hipEventSynchronize(s.seg_ev[2 * cells - 1]), "time cells sync");
    for (int i = 0; i < cells; ++i) {
        float ms = 0;
        OFP_HIP(c, hipEventElapsedTime(&ms, s.seg_ev[2 * i], s.seg_ev[2 * i + 1]), "event time");
        ms_out[i] = ms;
    }
    if (s.row >= 0) s.row += 1;
    return OFP_OK;
}

ofp_status ofp_set_pair_passes(ofp_ctx* c, int mode, uint32_t cell_mask) {
    if (!c) return OFP_E_ARG;
    // mode bits 8..9 select the pair-kernel variant (tuning)
    if (mode & ~0x3ff) return fail(c, OFP_E_ARG, "pair passes: unknown mode bits (group sweeps: ofp_set_group_launch)");
    const int variant = (mode >> 8) & 3;
    mode &= 0xff;
    if (mode < 0 || mode > kMaxKC || (mode > 1 && (mode & 1)))
        return fail(c, OFP_E_ARG, "pair passes: 0 (off), 1 (on) or an even moduli count per launch up to 16");
    if (variant != c->pair_variant) drop_graph(c->st);
    c->pair_variant = variant;
    const int on = mode > 0 ? 1 : 0;
    if (on != c->pair_rows) drop_graph(c->st);
    c->pair_rows = on;
    c->pair_kc = mode > 1 ? mode : kMaxKC;
    c->pair_mask = cell_mask;
    return OFP_OK;
}

ofp_status ofp_set_group_launch(ofp_ctx* c, int streams, long long pair_mask) {
    if (!c) return OFP_E_ARG;
    if (streams == 0) streams = 3;
    if (streams < 1 || streams > 4) return fail(c, OFP_E_ARG, "group streams: 1..4 (0 = the default, 3)");
    if (pair_mask < -1 || pair_mask >= (1ll << 32))
        return fail(c, OFP_E_ARG, "group pair mask: a 32-bit cell mask, or -1 for the library default");
    c->group_streams = streams;
    c->group_pair_mask = pair_mask;
    return OFP_OK;
}

ofp_status ofp_set_index_width(ofp_ctx* c, int bits) {
    if (!c) return OFP_E_ARG;
    if (bits != 0 && bits != 64) return fail(c, OFP_E_ARG, "index width must be 0 (by n) or 64");
    if (bits != c->index_bits) drop_graph(c->st);
    c->index_bits = bits;
    return OFP_OK;
}

ofp_status ofp_solve(ofp_ctx* c, int n, const uint32_t* moduli, int k, uint32_t* residues_out,
                     ofp_stats* st) {
    if (!c || !moduli || !residues_out) return OFP_E_ARG;
    auto t0 = std::chrono::steady_clock::now();
    uint64_t nm, nb;
    if (!counts(n, &nm, &nb)) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    if (k < 1 || k > 1024) return fail(c, OFP_E_ARG, "number of moduli must be in 1..1024");
    // Moduli per sweep: all of them when their interleaved state fits HBM (memory this
    // context already holds counts as free), else the largest count that fits, the
    // moduli then running in several sweeps one after another -- what the reference
    // does one modulus at a time (row6crt20.cu:1784).
    size_t fr = 0, tot = 0;
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    const uint64_t avail = fr + c->st.main_cap + c->st.blk_cap + c->win_bytes + c->stage_bytes;
    int g = c->max_batch > 0 ? std::min(k, c->max_batch) : k;
    while (g > 1 && state_bytes(c->layout, n, g) + c->reserve > avail) --g;
    if (state_bytes(c->layout, n, g) + c->reserve > avail)
        return fail(c, OFP_E_MEMORY, "insufficient HBM: one modulus needs " +
                                         std::to_string(state_bytes(c->layout, n, 1)) + " bytes, free " +
                                         std::to_string(avail));
    std::vector<int> sizes((size_t)k);
    const int sweeps = moduli_groups(k, g, sizes.data(), k);
    ofp_stats acc{};
    for (int i = 0, off = 0; i < sweeps; off += sizes[i], ++i) {
        ofp_status s = ensure_state(c, c->st, n, moduli + off, sizes[i]);
        if (s) return s;
        s = init_state(c, c->st);
        if (s) return s;
        ofp_stats rs{};
        s = ofp_row_sweep(c, c->st.W, &rs);
        if (s) return s;
        s = read_main(c, c->st, 0, residues_out + off, sizes[i]);  // rank_full(R) = 0 (fullmate_dropN_batch.cu:508,514)
        if (s) return s;
        if (i == 0) {
            acc = rs;
        } else {
            acc.kernel_s += rs.kernel_s;
            acc.transitions += rs.transitions;
            acc.rows += rs.rows;
            acc.launches += rs.launches;
            acc.full_kernel_s += rs.full_kernel_s;
            acc.capped_kernel_s += rs.capped_kernel_s;
            acc.full_launches += rs.full_launches;
            acc.capped_launches += rs.capped_launches;
        }
    }
    if (st) {
        *st = acc;
        st->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return OFP_OK;
}

ofp_status ofp_solve_static(ofp_ctx* c, int n, const uint32_t* moduli, int k, uint32_t* residues_out) {
    if (!c || !moduli || !residues_out) return OFP_E_ARG;
    uint64_t nm, nb;
    if (!counts(n, &nm, &nb)) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    if (k < 1 || k > kMaxKC) return fail(c, OFP_E_ARG, "ofp_solve_static takes 1..16 moduli");
    for (int j = 0; j < k; ++j)
        if (moduli[j] < 2) return fail(c, OFP_E_ARG, "modulus must be in [2, 4294967295]");
    const int Kp = interleave_stride(c->layout, k);
    const size_t mb = padded_states(nm) * (size_t)Kp * sizeof(uint32_t);
    const size_t bb = padded_states(nb ? nb : 1) * (size_t)Kp * sizeof(uint32_t);
    if (mb + bb > kStaticStateBytes)
        return fail(c, OFP_E_MEMORY, "state of " + std::to_string(mb + bb) + " bytes exceeds the " +
                                         std::to_string(kStaticStateBytes) + "-byte static array");
    if (!c->scratch) return fail(c, OFP_E_RUNTIME, "static array not set up");
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    Sweep& s = c->fixed;
    s.main = static_cast<uint32_t*>(c->scratch);
    s.blk = s.main + mb / sizeof(uint32_t);  // mb is a multiple of 256 bytes (64 states x 4 B)
    s.main_cap = mb;
    s.blk_cap = bb;
    s.mods_dev = nullptr;  // the start word takes the moduli by value
    s.mods.assign(moduli, moduli + k);
    s.n = n;
    s.W = n + 1;
    s.K = k;
    s.Kp = Kp;
    s.layout = c->layout;
    s.nmain = nm;
    s.nblk = nb;
    ofp_status st = init_state(c, s);  // clears both arrays: they hold the last call's state
    if (st) return st;
    int launches = 0;
    st = run_rows(c, s, s.W, &launches);
    if (st) return st;
    st = read_main(c, s, 0, residues_out, k);
    s.main = s.blk = nullptr;
    return st;
}

ofp_status ofp_set_max_batch(ofp_ctx* c, int max_moduli) {
    if (!c) return OFP_E_ARG;
    if (max_moduli < 0 || max_moduli > 1024) return fail(c, OFP_E_ARG, "max_moduli must be 0..1024");
    c->max_batch = max_moduli;
    return OFP_OK;
}

int ofp_moduli_groups(int k, int max_per_sweep, int* sizes, int capacity) {
    if (k < 1 || max_per_sweep < 1 || capacity < 0) return -1;
    return moduli_groups(k, max_per_sweep, sizes, capacity);
}

ofp_status ofp_apply_cells(ofp_ctx* c, int n, const uint32_t* moduli, int k, uint32_t* main_io,
                           uint32_t* blocked_io, int p_hi, int p_lo, int blocked_valid) {
    if (!c || !moduli || !main_io || !blocked_io) return OFP_E_ARG;
    if (n < 1 || n > OFP_MAX_N) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    const int W = n + 1;
    if (p_hi > W - 1 || p_lo < 1 || p_hi < p_lo) return fail(c, OFP_E_ARG, "need W-1 >= p_hi >= p_lo >= 1");
    Sweep& sw = c->st;
    ofp_status s = ensure_state(c, sw, n, moduli, k);
    const bool pairs = pairs_apply(c, n) && p_hi - p_lo >= 1;
    if (s) return s;
    // host state-major (index*k + j) <-> device layout
    const int Kp = sw.Kp, lay = sw.layout;
    std::vector<uint32_t> tm(padded_states(sw.nmain) * (size_t)Kp, 0),
        tb(padded_states(sw.nblk ? sw.nblk : 1) * (size_t)Kp, 0);
    for (uint64_t i = 0; i < sw.nmain; ++i)
        for (int j = 0; j < k; ++j) tm[state_elem(lay, i, Kp, j)] = main_io[i * k + j];
    for (uint64_t i = 0; i < sw.nblk; ++i)
        for (int j = 0; j < k; ++j) tb[state_elem(lay, i, Kp, j)] = blocked_io[i * k + j];
    OFP_HIP(c, hipMemcpy(sw.main, tm.data(), tm.size() * 4, hipMemcpyHostToDevice), "upload main");
    OFP_HIP(c, hipMemcpy(sw.blk, tb.data(), tb.size() * 4, hipMemcpyHostToDevice), "upload blocked");
    sw.row = -1;  // arbitrary state: no height caps
    s = pairs ? enqueue_pairs(c, sw, p_hi, p_lo, blocked_valid != 0, nullptr)
              : enqueue_cells(c, sw, p_hi, p_lo, blocked_valid != 0, nullptr);
    if (s) return s;
    OFP_HIP(c, hipStreamSynchronize(sw.stream), "apply sync");
    OFP_HIP(c, hipMemcpy(tm.data(), sw.main, tm.size() * 4, hipMemcpyDeviceToHost), "download main");
    OFP_HIP(c, hipMemcpy(tb.data(), sw.blk, tb.size() * 4, hipMemcpyDeviceToHost), "download blocked");
    for (uint64_t i = 0; i < sw.nmain; ++i)
        for (int j = 0; j < k; ++j) main_io[i * k + j] = tm[state_elem(lay, i, Kp, j)];
    // after cell 1 nothing is blocked; the device buffer is simply not written
    for (uint64_t i = 0; i < sw.nblk; ++i)
        for (int j = 0; j < k; ++j) blocked_io[i * k + j] = (p_lo == 1) ? 0u : tb[state_elem(lay, i, Kp, j)];
    return OFP_OK;
}

// ---- shard sweeps: one GPU's share of the low cells of a row ----

ofp_status ofp_shard_plan(ofp_ctx* c, const ofp_shard_spec* sp, ofp_shard_info* info) {
    // host-only computation: c may be NULL (then no error message is kept)
    if (!sp || !info) return OFP_E_ARG;
    uint64_t nm, nb;
    if (!counts(sp->n, &nm, &nb)) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    const int W = sp->n + 1;
    const int lc = sp->low_cells ? sp->low_cells : W / 2;
    if (lc < 1 || lc > W - 2) return fail(c, OFP_E_ARG, "low_cells must be in 1..W-2");
    if (sp->slices < 1 || sp->slice < 0 || sp->slice >= sp->slices) return fail(c, OFP_E_ARG, "bad slice");
    if (sp->k < 1 || sp->k > 8 || (sp->k & (sp->k - 1))) return fail(c, OFP_E_ARG, "shard k must be 1, 2, 4 or 8");
    const uint64_t r0 = (uint64_t)((unsigned __int128)nb * (unsigned)sp->slice / (unsigned)sp->slices);
    const uint64_t r1 = (uint64_t)((unsigned __int128)nb * (unsigned)(sp->slice + 1) / (unsigned)sp->slices);
    uint64_t b0, m0, b1 = nb, m1 = nm;
    class_start(W, lc, r0, &b0, &m0);
    if (r1 < nb) class_start(W, lc, r1, &b1, &m1);
    info->first_blocked = b0;
    info->blocked_count = b1 - b0;
    info->first_main = m0;
    info->main_count = m1 - m0;
    info->low_cells = lc;
    const int Kp = interleave_stride(1, sp->k);
    info->state_bytes = (info->main_count + info->blocked_count) * (uint64_t)Kp * 4;
    info->kernel_s = 0;
    info->launches = 0;
    info->stray_stores = 0;
    return OFP_OK;
}

ofp_status ofp_shard_sweep(ofp_ctx* c, const ofp_shard_spec* sp, ofp_shard_info* info, uint32_t* main_out,
                           uint32_t* blk_out) {
    if (!c) return OFP_E_ARG;
    ofp_status st = ofp_shard_plan(c, sp, info);
    if (st) return st;
    for (int j = 0; j < sp->k; ++j)
        if (sp->moduli[j] < 2) return fail(c, OFP_E_ARG, "modulus must be in [2, 4294967295]");
    const int W = sp->n + 1, k = sp->k, Kp = interleave_stride(1, k);
    const uint64_t M = info->main_count, B = info->blocked_count;
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    const size_t mbytes = std::max<uint64_t>(M, 1) * Kp * 4, bbytes = std::max<uint64_t>(B, 1) * Kp * 4;
    char* wbuf = nullptr;
    c->last_low.valid = false;
    st = window_buffer(c, (mbytes + 255) / 256 * 256 + (bbytes + 255) / 256 * 256 + 512, &wbuf);
    if (st) return st;
    Carve cv{wbuf};
    uint32_t* dm = cv.take<uint32_t>(mbytes);
    uint32_t* db = cv.take<uint32_t>(bbytes);
    uint32_t* dmods = cv.take<uint32_t>(16 * 4);
    unsigned long long* dstray = cv.take<unsigned long long>(sizeof(unsigned long long));
    auto release = [&]() {};
#define OFP_HIP_R(expr, what)                                        \
    do {                                                             \
        hipError_t e_ = (expr);                                      \
        if (e_ != hipSuccess) { release(); return hip_fail(c, e_, what); } \
    } while (0)
    OFP_HIP_R(hipMemcpy(dmods, sp->moduli, k * 4, hipMemcpyHostToDevice), "copy moduli");
    OFP_HIP_R(hipMemset(dstray, 0, sizeof(unsigned long long)), "zero counter");
    hipError_t e = launch_fill_remote(dm, info->first_main, M, db, info->first_blocked, B, Kp, k, dmods,
                                      sp->seed, c->stream);
    if (e != hipSuccess) { release(); return hip_fail(c, e, "fill shard window"); }
    const bool wide = W > 23 || c->index_bits == 64;
    const int blocks = grid_blocks(c, B);
    OFP_HIP_R(hipEventRecord(c->ev0, c->stream), "event");
    int launches = 0;
    for (int p = info->low_cells; p >= 1; --p) {
        for (int off = 0; off < k; off += 8) {
            CellLaunch a{};
            a.main = dm;
            a.blk = db;
            a.dp = wide ? c->dp64 : c->dp32;
            a.nblk = B;
            a.first = info->first_blocked;
            a.layout = 1;
            a.variant = c->variant;  // shard cells are never sorted (slower at 64-bit ranks)
            a.xcd_shift = piece_shift(c->xcd_shift, B, c->threads);
            a.W = W;
            a.p = p;
            a.read_blk = (p == W - 1) ? 0 : 1;
            a.K = Kp;
            a.off = off;
            a.kc = std::min(8, k - off);
            for (int j = 0; j < a.kc; ++j) a.mod[j] = sp->moduli[off + j];
            a.m0 = info->first_main;
            a.M = M;
            a.seed = sp->seed;
            a.stray = dstray;
            e = launch_cell_shard(a, wide, c->threads, blocks, c->stream);
            if (e != hipSuccess) { release(); return hip_fail(c, e, "shard cell launch"); }
            ++launches;
        }
    }
    OFP_HIP_R(hipEventRecord(c->ev1, c->stream), "event");
    OFP_HIP_R(hipEventSynchronize(c->ev1), "shard sync");
    float ms = 0;
    OFP_HIP_R(hipEventElapsedTime(&ms, c->ev0, c->ev1), "event time");
    info->kernel_s = ms * 1e-3;
    info->launches = launches;
    unsigned long long stray = 0;
    OFP_HIP_R(hipMemcpy(&stray, dstray, sizeof stray, hipMemcpyDeviceToHost), "read counter");
    info->stray_stores = stray;
    if (main_out || blk_out) {
        std::vector<uint32_t> t(std::max(M, B) * (size_t)Kp);
        if (main_out && M) {
            OFP_HIP_R(hipMemcpy(t.data(), dm, M * Kp * 4, hipMemcpyDeviceToHost), "download main");
            for (uint64_t i = 0; i < M; ++i)
                for (int j = 0; j < k; ++j) main_out[i * k + j] = t[i * Kp + j];
        }
        if (blk_out && B) {
            OFP_HIP_R(hipMemcpy(t.data(), db, B * Kp * 4, hipMemcpyDeviceToHost), "download blocked");
            for (uint64_t i = 0; i < B; ++i)
                for (int j = 0; j < k; ++j) blk_out[i * k + j] = t[i * Kp + j];
        }
    }
#undef OFP_HIP_R
    release();
    return OFP_OK;
}

// ---- group sweeps: the high cells of a row on low-occupancy groups ----

namespace {

ofp_status group_check(ofp_ctx* c, const ofp_group_spec* sp, int* low) {
    if (!sp || !sp->moduli) return fail(c, OFP_E_ARG, "group spec and moduli are required");
    if (sp->n < 2 || sp->n > OFP_MAX_N) return fail(c, OFP_E_RANGE, "group sweeps take n in 2..27");
    const int W = sp->n + 1;
    const int L = sp->low ? sp->low : W / 2;
    if (sp->high != 0 && sp->high != 1) return fail(c, OFP_E_ARG, "high must be 0 or 1");
    const int fixed = sp->high ? W - 1 - L : L;
    if (L < 1 || L > W - 2 || fixed > 24)
        return fail(c, OFP_E_ARG, "low must be in 1..W-2 with at most 24 fixed positions");
    if (sp->k < 1 || sp->k > kMaxKC) return fail(c, OFP_E_ARG, "group sweeps take 1..16 moduli");
    const uint64_t all = 1ull << fixed;
    if (sp->groups < 1 || sp->first_group >= all || sp->groups > all - sp->first_group)
        return fail(c, OFP_E_ARG, "groups must lie in 0..2^low-1");
    *low = L;
    return OFP_OK;
}

constexpr size_t kGroupTableBytes = sizeof(cell::Quad<uint32_t>) * kDpRows * kDpCols;

}  // namespace

ofp_status ofp_group_plan(ofp_ctx* c, const ofp_group_spec* sp, ofp_group_info* info, uint64_t* main_counts,
                          uint64_t* blocked_counts) {
    if (!info) return fail(c, OFP_E_ARG, "info is required");
    int L = 0;
    ofp_status st = group_check(c, sp, &L);
    if (st) return st;
    const int W = sp->n + 1;
    uint64_t M = 0, B = 0, big = 0;
    for (uint32_t g = 0; g < sp->groups; ++g) {
        uint64_t m, b;
        group_counts(sp->high != 0, W, L, sp->first_group + g, &m, &b);
        if (main_counts) main_counts[g] = m;
        if (blocked_counts) blocked_counts[g] = b;
        M += m;
        B += b;
        big = std::max(big, m + b);
    }
    *info = ofp_group_info{};
    info->main_count = M;
    info->blocked_count = B;
    info->largest_group = big;
    const int Kp = interleave_stride(1, sp->k);
    info->state_bytes = (M + B) * (uint64_t)Kp * 4 + (uint64_t)sp->groups * kGroupTableBytes;
    info->low = L;
    info->cells = sp->high ? L : W - 1 - L;
    return OFP_OK;
}

ofp_status ofp_group_sweep(ofp_ctx* c, const ofp_group_spec* sp, ofp_group_info* info, uint32_t* main_io,
                           uint32_t* blk_io) {
    if (!c || !sp || !info) return fail(c, OFP_E_ARG, "ctx, spec and info are required");
    const uint32_t G = sp->groups;
    std::vector<uint64_t> mc(G), bc(G);
    ofp_status st = ofp_group_plan(c, sp, info, mc.data(), bc.data());
    if (st) return st;
    for (int j = 0; j < sp->k; ++j)
        if (sp->moduli[j] < 2) return fail(c, OFP_E_ARG, "modulus must be in [2, 4294967295]");
    for (uint32_t g = 0; g < G; ++g)
        if ((mc[g] >> 32) || (bc[g] >> 32))
            return fail(c, OFP_E_ARG, "a group of the window holds 2^32 or more words: use more low positions");
    const int W = sp->n + 1, k = sp->k, Kp = interleave_stride(1, k), L = info->low;
    const uint64_t M = info->main_count, B = info->blocked_count;
    std::vector<uint64_t> moff(G + 1, 0), boff(G + 1, 0);
    for (uint32_t g = 0; g < G; ++g) {
        moff[g + 1] = moff[g] + mc[g];
        boff[g + 1] = boff[g] + bc[g];
    }
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    const size_t mbytes = std::max<uint64_t>(M, 1) * Kp * 4, bbytes = std::max<uint64_t>(B, 1) * Kp * 4;
    const size_t tbytes = (size_t)G * kGroupTableBytes;
    // The half's steps: cells p and p-1 as one pair pass where the group pair
    // mask (ofp_set_group_launch) pairs them, else one cell.
    const int p_hi = sp->high ? L : W - 1, p_lo = sp->high ? 1 : L + 1;
    const uint32_t pmask = c->group_pair_mask >= 0 ? (uint32_t)c->group_pair_mask : default_group_pair_mask();
    std::vector<int> steps;  // p of each step; a pair step is marked by p | 0x100
    for (int p = p_hi; p >= p_lo;) {
        if (p - 1 >= p_lo && p >= 2 && ((pmask >> p) & 1u)) {
            steps.push_back(p | 0x100);
            p -= 2;
        } else {
            steps.push_back(p);
            --p;
        }
    }
    int npairs = 0;
    for (int x : steps) npairs += (x & 0x100) ? 1 : 0;
    // Per group, pair step and window change d (-1, 0, 1, -2, 2): the group's
    // class table and class count.
    static const int kPairD[5] = {-1, 0, 1, -2, 2};
    const size_t ctb = class_table_bytes(false);
    const size_t ctbytes = (size_t)G * npairs * 5 * ctb;
    std::vector<char> cts(std::max<size_t>(ctbytes, 1));
    std::vector<uint64_t> ccount((size_t)G * npairs * 5, 0);
    if (npairs) {
        static uint64_t fm[kDpRows][kDpCols];
        static cell::Pair<uint32_t> t[kDpRows][kDpCols];
        for (uint32_t g = 0; g < G; ++g) {
            cell::Occupancy om, ob;
            cell::group_occupancy(sp->high != 0, W, L, sp->first_group + g, om, ob);
            cell::build_group_table(om, fm);
            int j = 0;
            for (int x : steps) {
                if (!(x & 0x100)) continue;
                const int p = x & 0xff;
                for (int di = 0; di < 5; ++di) {
                    const size_t slot = ((size_t)g * npairs + j) * 5 + di;
                    pair::pack_group_class_table<uint32_t, kDpCols>(om, fm, p, kPairD[di], t);
                    std::memcpy(cts.data() + slot * ctb, t, ctb);
                    ccount[slot] = pair::group_class_count(om, fm, W, p, kPairD[di]);
                }
                ++j;
            }
        }
    }
    char* wbuf = nullptr;
    c->last_low.valid = false;
    st = window_buffer(c, (mbytes + 255) / 256 * 256 + (bbytes + 255) / 256 * 256 + tbytes + ctbytes + 512, &wbuf);
    if (st) return st;
    Carve cv{wbuf};
    uint32_t* dm = cv.take<uint32_t>(mbytes);
    uint32_t* db = cv.take<uint32_t>(bbytes);
    char* dtab = cv.take<char>(tbytes);
    char* dct = cv.take<char>(ctbytes);
    uint32_t* dmods = cv.take<uint32_t>(kMaxKC * 4);
    std::vector<hipEvent_t> ev(info->cells + 1, nullptr);
    auto release = [&]() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
    };
#define OFP_HIP_R(expr, what)                                              \
    do {                                                                   \
        hipError_t e_ = (expr);                                            \
        if (e_ != hipSuccess) { release(); return hip_fail(c, e_, what); } \
    } while (0)
    for (hipEvent_t& e : ev) OFP_HIP_R(hipEventCreate(&e), "group event");
    {
        std::vector<char> tabs((size_t)G * kGroupTableBytes);
        for (uint32_t g = 0; g < G; ++g)
            if (pack_group_table(sp->high != 0, W, L, sp->first_group + g,
                                 tabs.data() + (size_t)g * kGroupTableBytes, kGroupTableBytes) != kGroupTableBytes) {
                release();
                return fail(c, OFP_E_RUNTIME, "group table packing");
            }
        OFP_HIP_R(hipMemcpy(dtab, tabs.data(), tabs.size(), hipMemcpyHostToDevice), "upload group tables");
        if (ctbytes) OFP_HIP_R(hipMemcpy(dct, cts.data(), ctbytes, hipMemcpyHostToDevice), "upload class tables");
    }
    OFP_HIP_R(hipMemcpy(dmods, sp->moduli, k * 4, hipMemcpyHostToDevice), "copy moduli");
    const bool blk_in = sp->high && blk_io;
    if (main_io || blk_in) {
        std::vector<uint32_t> t(std::max<uint64_t>(std::max(M, B), 1) * Kp, 0);
        if (main_io) {
            for (uint64_t i = 0; i < M; ++i)
                for (int j = 0; j < k; ++j) t[i * Kp + j] = main_io[i * k + j];
            OFP_HIP_R(hipMemcpy(dm, t.data(), M * Kp * 4, hipMemcpyHostToDevice), "upload main window");
        }
        if (blk_in) {
            for (uint64_t i = 0; i < B; ++i)
                for (int j = 0; j < k; ++j) t[i * Kp + j] = blk_io[i * k + j];
            OFP_HIP_R(hipMemcpy(db, t.data(), B * Kp * 4, hipMemcpyHostToDevice), "upload blocked window");
        }
    }
    // Seeded inputs, one hash domain per window start and half: main words
    // unless given, blocked words only where the first cell reads them.
    const uint64_t dom = ((uint64_t)sp->first_group << 36) | ((uint64_t)sp->high << 35);
    const uint64_t fillM = main_io ? 0 : M, fillB = (sp->high && !blk_io) ? B : 0;
    if (fillM || fillB)
        OFP_HIP_R(launch_fill_remote(dm, dom, fillM, db, dom, fillB, Kp, k, dmods, sp->seed, c->stream),
                  "fill group window");
    const int sf = c->sorted_from >= 0 ? c->sorted_from : (k >= 8 ? 6 : 0);
    if (c->group_streams > 1 && !c->gfork) {  // gfork is created last: all of them exist
        for (int q = 0; q < 3; ++q) {
            if (!c->gstream[q])
                OFP_HIP_R(hipStreamCreateWithFlags(&c->gstream[q], hipStreamNonBlocking), "group stream");
            if (!c->gjoin[q]) OFP_HIP_R(hipEventCreateWithFlags(&c->gjoin[q], hipEventDisableTiming), "group event");
        }
        OFP_HIP_R(hipEventCreateWithFlags(&c->gfork, hipEventDisableTiming), "group event");
    }
    // The groups of a step are independent (each has its own words), so their
    // launches are dealt to ns streams that run side by side: the many small
    // groups' launches fill each other's ramp-up and tail.  The streams fork
    // from and join the sweep's stream at every step (a cell or a pair pass),
    // which keeps the step boundaries (and per-step times) of the one-stream
    // order; a group keeps one stream within a step.
    const int ns = c->group_streams;
    auto fork = [&]() -> ofp_status {
        if (ns > 1) {
            OFP_HIP_R(hipEventRecord(c->gfork, c->stream), "group fork");
            for (int q = 0; q < ns - 1; ++q) OFP_HIP_R(hipStreamWaitEvent(c->gstream[q], c->gfork, 0), "group fork");
        }
        return OFP_OK;
    };
    auto join = [&]() -> ofp_status {
        if (ns > 1)
            for (int q = 0; q < ns - 1; ++q) {
                OFP_HIP_R(hipEventRecord(c->gjoin[q], c->gstream[q]), "group join");
                OFP_HIP_R(hipStreamWaitEvent(c->stream, c->gjoin[q], 0), "group join");
            }
        return OFP_OK;
    };
    // Groups go to the stream with the least work so far (blocked words), in
    // group order: group sizes follow the mask bits, so a round-robin deal
    // would give one stream most of the large groups.
    std::vector<uint8_t> gsi(G, 0);
    auto deal = [&]() {
        uint64_t load[4] = {0, 0, 0, 0};
        for (uint32_t g = 0; g < G; ++g) {
            int q = 0;
            for (int r = 1; r < ns; ++r)
                if (load[r] < load[q]) q = r;
            gsi[g] = (uint8_t)q;
            load[q] += bc[g] + 1;
        }
    };
    deal();
    auto stream_of = [&](uint32_t g) { return gsi[g] == 0 ? c->stream : c->gstream[gsi[g] - 1]; };
    int launches = 0;
    OFP_HIP_R(hipEventRecord(ev[0], c->stream), "event");
    int i = 0, pj = 0;
    for (int x : steps) {
        const int p = x & 0xff;
        if (x & 0x100) {
            // pair pass: per window change d, every group (inside a group the
            // launches keep the order d = -1, 0, 1, -2, 2 on the group's stream)
            if (ofp_status f = fork()) return f;
            for (int di = 0; di < 5; ++di)
                for (uint32_t g = 0; g < G; ++g) {
                    const size_t slot = ((size_t)g * npairs + pj) * 5 + di;
                    const uint64_t cnt = ccount[slot];
                    if (!bc[g] || !cnt) continue;
                    pair::PairLaunch a{};
                    a.main_in = dm + moff[g] * Kp;
                    a.blk_in = db + boff[g] * Kp;
                    a.main_out = dm + moff[g] * Kp;
                    a.blk_out = db + boff[g] * Kp;
                    a.dp = dtab + (size_t)g * kGroupTableBytes;
                    a.ct = dct + slot * ctb;
                    a.classes = cnt;
                    a.W = W;
                    a.p = p;
                    a.d = kPairD[di];
                    a.read_blk = (p == W - 1) ? 0 : 1;
                    a.K = Kp;
                    a.off = 0;
                    a.kc = Kp;
                    for (int jj = 0; jj < Kp; ++jj) a.mod[jj] = sp->moduli[jj < k ? jj : 0];
                    a.xcd_shift = piece_shift(c->xcd_shift, cnt, c->threads);
                    a.variant = c->pair_variant;
                    OFP_HIP_R(launch_pair(a, false, c->threads, grid_blocks(c, cnt), stream_of(g)), "group pair launch");
                    ++launches;
                }
            if (ofp_status f = join()) return f;
            OFP_HIP_R(hipEventRecord(ev[i + 1], c->stream), "event");
            OFP_HIP_R(hipEventRecord(ev[i + 2], c->stream), "event");
            i += 2;
            ++pj;
            continue;
        }
        if (ofp_status f = fork()) return f;
        for (uint32_t g = 0; g < G; ++g) {
            if (!bc[g]) continue;  // no blocked word: the cell changes nothing in the group
            hipStream_t gs = stream_of(g);
            CellLaunch a{};
            a.main = dm + moff[g] * Kp;
            a.blk = db + boff[g] * Kp;
            a.dp = dtab + (size_t)g * kGroupTableBytes;
            a.nblk = bc[g];
            a.first = 0;
            a.layout = 1;
            a.variant = (sf > 0 && p >= sf) ? 4 : c->variant;
            a.xcd_shift = piece_shift(c->xcd_shift, bc[g], c->threads);
            a.W = W;
            a.p = p;
            a.read_blk = (p == W - 1) ? 0 : 1;
            a.K = Kp;
            a.off = 0;
            a.kc = k;
            for (int j = 0; j < k; ++j) a.mod[j] = sp->moduli[j];
            OFP_HIP_R(launch_cell(a, false, c->threads, grid_blocks(c, bc[g]), gs), "group cell launch");
            ++launches;
        }
        if (ofp_status f = join()) return f;
        OFP_HIP_R(hipEventRecord(ev[i + 1], c->stream), "event");
        ++i;
    }
    OFP_HIP_R(hipEventSynchronize(ev[info->cells]), "group sync");
    double total = 0;
    for (int i = 0; i < info->cells; ++i) {
        float ms = 0;
        OFP_HIP_R(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "event time");
        if (i < OFP_MAX_N) info->cell_ms[i] = ms;
        total += ms;
    }
    info->kernel_s = total * 1e-3;
    info->launches = launches;
    info->steps = (int)steps.size();
    if (!sp->high) {
        ofp_ctx::LowWindow& lw = c->last_low;
        lw.n = sp->n;
        lw.k = k;
        lw.low = L;
        lw.first = sp->first_group;
        lw.groups = G;
        lw.main = dm;
        lw.blk = db;
        lw.tables = dtab;
        lw.mc = mc;
        lw.bc = bc;
        lw.valid = true;
    }
    const bool blk_out = !sp->high && blk_io;
    if (main_io || blk_out) {
        std::vector<uint32_t> t(std::max<uint64_t>(std::max(M, B), 1) * Kp);
        if (main_io && M) {
            OFP_HIP_R(hipMemcpy(t.data(), dm, M * Kp * 4, hipMemcpyDeviceToHost), "download main window");
            for (uint64_t i = 0; i < M; ++i)
                for (int j = 0; j < k; ++j) main_io[i * k + j] = t[i * Kp + j];
        }
        if (blk_out && B) {
            OFP_HIP_R(hipMemcpy(t.data(), db, B * Kp * 4, hipMemcpyDeviceToHost), "download blocked window");
            for (uint64_t i = 0; i < B; ++i)
                for (int j = 0; j < k; ++j) blk_io[i * k + j] = t[i * Kp + j];
        }
    }
#undef OFP_HIP_R
    release();
    return OFP_OK;
}

// ---- exchange: the words of a LOW window to their HIGH groups' shares ----

namespace {

// The exchange buffer: grown on demand, kept from call to call like the window.
ofp_status stage_buffer(ofp_ctx* c, size_t bytes, char** out) {
    bytes = (bytes + 255) / 256 * 256;
    if (c->stage && c->stage_bytes >= bytes) {
        *out = c->stage;
        return OFP_OK;
    }
    if (c->stage) {
        OFP_HIP(c, hipFree(c->stage), "free exchange buffer");
        c->stage = nullptr;
        c->stage_bytes = 0;
    }
    size_t fr = 0, tot = 0;
    OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    if (bytes + c->reserve > fr) return fail(c, OFP_E_MEMORY, "exchange staging does not fit HBM beside the window");
    const size_t slack = bytes + c->reserve + bytes / 16 <= fr ? bytes / 16 / 256 * 256 : 0;  // as the window's
    OFP_HIP(c, hipMalloc(&c->stage, bytes + slack), "exchange buffer");
    c->stage_bytes = bytes + slack;
    *out = c->stage;
    return OFP_OK;
}

// Tiles of one kind of word of a window: every group cut into runs of at most
// kXTileWords words.
void window_tiles(const std::vector<uint64_t>& counts, std::vector<XTile>& t) {
    t.clear();
    uint64_t word0 = 0;
    for (uint32_t g = 0; g < counts.size(); ++g) {
        for (uint64_t r = 0; r < counts[g]; r += kXTileWords) {
            XTile x{};
            x.word0 = word0 + r;
            x.group = g;
            x.rank0 = (uint32_t)r;
            x.len = (uint32_t)std::min<uint64_t>(kXTileWords, counts[g] - r);
            t.push_back(x);
        }
        word0 += counts[g];
    }
}

}  // namespace

ofp_status ofp_group_exchange(ofp_ctx* c, const ofp_group_spec* sp, int shares, ofp_exchange_info* info,
                              uint64_t* main_index, uint32_t* main_res, uint64_t* blocked_index,
                              uint32_t* blocked_res) {
    if (!c || !sp || !info) return fail(c, OFP_E_ARG, "ctx, spec and info are required");
    if (shares < 1 || shares > kMaxShares) return fail(c, OFP_E_ARG, "shares must be in 1..16");
    int L = 0;
    ofp_status st = group_check(c, sp, &L);
    if (st) return st;
    if (sp->high) return fail(c, OFP_E_ARG, "the exchange takes a LOW window (high = 0)");
    const ofp_ctx::LowWindow& lw = c->last_low;
    if (!lw.valid || lw.n != sp->n || lw.k != sp->k || lw.low != L || lw.first != sp->first_group ||
        lw.groups != sp->groups)
        return fail(c, OFP_E_ARG, "no such LOW window in the context: sweep it with ofp_group_sweep first");
    const int W = sp->n + 1, k = sp->k, Kp = interleave_stride(1, k), nf = W - 1 - L;
    if (nf > 16) return fail(c, OFP_E_ARG, "the exchange takes at most 16 fixed HIGH positions");
    const uint32_t NM = 1u << nf;
    // HIGH groups: word counts, the shares (consecutive masks, cut by main +
    // blocked words at s/shares of the total: the split ofp.shard uses) and
    // each group's offset in its share's placement.
    std::vector<uint64_t> hm(NM), hb(NM);
    for (uint32_t m = 0; m < NM; ++m) {
        group_counts(true, W, L, m, &hm[m], &hb[m]);
        if ((hm[m] >> 32) || (hb[m] >> 32))
            return fail(c, OFP_E_ARG, "a HIGH group of the split holds 2^32 or more words: use a lower split");
    }
    std::vector<double> cum(NM + 1, 0.0);
    for (uint32_t m = 0; m < NM; ++m) cum[m + 1] = cum[m] + (double)hm[m] + (double)hb[m];
    uint32_t first[kMaxShares + 1] = {};
    for (int s = 0; s < shares; ++s) {
        const double x = cum[NM] * s / shares;
        first[s] = (uint32_t)(std::lower_bound(cum.begin(), cum.end(), x) - cum.begin());
    }
    first[shares] = NM;
    std::vector<uint64_t> base_m(NM), base_b(NM);
    for (int s = 0; s < shares; ++s) {
        uint64_t rm = 0, rb = 0;
        for (uint32_t m = first[s]; m < first[s + 1]; ++m) {
            base_m[m] = rm;
            base_b[m] = rb;
            rm += hm[m];
            rb += hb[m];
        }
    }
    uint64_t d[kDpRows][kDpCols];
    build_height_table(d);
    const size_t rows = (size_t)(NM - 1) * kDpCols;
    std::vector<uint32_t> ht_m(std::max<size_t>(rows, 1)), ht_b(std::max<size_t>(rows, 1));
    xchg::build_high_rows(d, L + 1, nf, ht_m.data());
    xchg::build_high_rows(d, L, nf, ht_b.data());
    std::vector<XTile> tm, tb;
    window_tiles(lw.mc, tm);
    window_tiles(lw.bc, tb);
    // the window's groups' bottom tables: main words below position low+1,
    // blocked words below position low
    std::vector<uint32_t> lut_m, lut_b, off_m((size_t)lw.groups * (kDpCols + 1)),
        off_b((size_t)lw.groups * (kDpCols + 1));
    {
        std::vector<uint32_t> part;
        uint64_t f[kDpRows][kDpCols];
        for (uint32_t g = 0; g < lw.groups; ++g) {
            cell::Occupancy om, ob;
            cell::group_occupancy(false, W, L, lw.first + g, om, ob);
            cell::build_group_table(om, f);
            for (int kind = 0; kind < 2; ++kind) {
                std::vector<uint32_t>& lut = kind == 0 ? lut_m : lut_b;
                uint32_t* off = (kind == 0 ? off_m : off_b).data() + (size_t)g * (kDpCols + 1);
                xchg::build_bottom_lut(om, f, d, kind == 0 ? L + 1 : L, part, off);
                const uint32_t at = (uint32_t)lut.size();
                for (int h = 0; h <= kDpCols; ++h) off[h] += at;
                lut.insert(lut.end(), part.begin(), part.end());
            }
        }
    }
    uint64_t M = 0, B = 0;
    for (uint64_t x : lw.mc) M += x;
    for (uint64_t x : lw.bc) B += x;

    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_ht = al(ht_m.size() * 4), b_base = al((size_t)NM * 8);
    const size_t b_tm = al(std::max<size_t>(tm.size(), 1) * sizeof(XTile)),
                 b_tb = al(std::max<size_t>(tb.size(), 1) * sizeof(XTile));
    const size_t b_cm = al(std::max<size_t>(tm.size(), 1) * shares * 4),
                 b_cb = al(std::max<size_t>(tb.size(), 1) * shares * 4);
    const size_t b_om = al(std::max<size_t>(tm.size(), 1) * shares * 8),
                 b_ob = al(std::max<size_t>(tb.size(), 1) * shares * 8);
    const size_t b_tot = al(2 * kMaxShares * 8);
    const size_t b_lm = al(std::max<size_t>(lut_m.size(), 1) * 4), b_lb = al(std::max<size_t>(lut_b.size(), 1) * 4);
    const size_t b_off = al(off_m.size() * 4);
    const size_t b_im = al(std::max<uint64_t>(M, 1) * 8), b_ib = al(std::max<uint64_t>(B, 1) * 8);
    const size_t b_rm = al(std::max<uint64_t>(M, 1) * Kp * 4), b_rb = al(std::max<uint64_t>(B, 1) * Kp * 4);
    const size_t need = 2 * b_ht + 2 * b_base + b_tm + b_tb + b_cm + b_cb + b_om + b_ob + b_tot + b_lm + b_lb +
                        2 * b_off + b_im + b_ib + b_rm + b_rb;
    char* buf = nullptr;
    st = stage_buffer(c, need, &buf);
    if (st) return st;
    Carve cv{buf};
    uint32_t* d_htm = cv.take<uint32_t>(b_ht);
    uint32_t* d_htb = cv.take<uint32_t>(b_ht);
    uint64_t* d_bm = cv.take<uint64_t>(b_base);
    uint64_t* d_bb = cv.take<uint64_t>(b_base);
    XTile* d_tm = cv.take<XTile>(b_tm);
    XTile* d_tb = cv.take<XTile>(b_tb);
    uint32_t* d_cm = cv.take<uint32_t>(b_cm);
    uint32_t* d_cb = cv.take<uint32_t>(b_cb);
    uint64_t* d_om = cv.take<uint64_t>(b_om);
    uint64_t* d_ob = cv.take<uint64_t>(b_ob);
    uint64_t* d_tot = cv.take<uint64_t>(b_tot);
    uint32_t* d_lm = cv.take<uint32_t>(b_lm);
    uint32_t* d_lb = cv.take<uint32_t>(b_lb);
    uint32_t* d_offm = cv.take<uint32_t>(b_off);
    uint32_t* d_offb = cv.take<uint32_t>(b_off);
    uint64_t* d_im = cv.take<uint64_t>(b_im);
    uint64_t* d_ib = cv.take<uint64_t>(b_ib);
    uint32_t* d_rm = cv.take<uint32_t>(b_rm);
    uint32_t* d_rb = cv.take<uint32_t>(b_rb);
    OFP_HIP(c, hipMemcpy(d_htm, ht_m.data(), ht_m.size() * 4, hipMemcpyHostToDevice), "upload group rows");
    OFP_HIP(c, hipMemcpy(d_htb, ht_b.data(), ht_b.size() * 4, hipMemcpyHostToDevice), "upload group rows");
    OFP_HIP(c, hipMemcpy(d_bm, base_m.data(), (size_t)NM * 8, hipMemcpyHostToDevice), "upload bases");
    OFP_HIP(c, hipMemcpy(d_bb, base_b.data(), (size_t)NM * 8, hipMemcpyHostToDevice), "upload bases");
    if (!lut_m.empty()) OFP_HIP(c, hipMemcpy(d_lm, lut_m.data(), lut_m.size() * 4, hipMemcpyHostToDevice), "upload bottoms");
    if (!lut_b.empty()) OFP_HIP(c, hipMemcpy(d_lb, lut_b.data(), lut_b.size() * 4, hipMemcpyHostToDevice), "upload bottoms");
    OFP_HIP(c, hipMemcpy(d_offm, off_m.data(), off_m.size() * 4, hipMemcpyHostToDevice), "upload bottom offsets");
    OFP_HIP(c, hipMemcpy(d_offb, off_b.data(), off_b.size() * 4, hipMemcpyHostToDevice), "upload bottom offsets");
    if (!tm.empty())
        OFP_HIP(c, hipMemcpy(d_tm, tm.data(), tm.size() * sizeof(XTile), hipMemcpyHostToDevice), "upload tiles");
    if (!tb.empty())
        OFP_HIP(c, hipMemcpy(d_tb, tb.data(), tb.size() * sizeof(XTile), hipMemcpyHostToDevice), "upload tiles");

    auto launch_of = [&](bool main_kind) {
        XchgLaunch a{};
        a.res = main_kind ? lw.main : lw.blk;
        a.gtab = lw.tables;
        a.tiles = main_kind ? d_tm : d_tb;
        a.ntiles = (uint32_t)(main_kind ? tm.size() : tb.size());
        a.K = Kp;
        a.top = main_kind ? W - 1 : W - 2;
        a.base = main_kind ? L + 1 : L;
        a.nf = nf;
        a.main = main_kind ? 1 : 0;
        a.ht = main_kind ? d_htm : d_htb;
        a.dplain = c->dp32;
        a.base_of = main_kind ? d_bm : d_bb;
        a.lut = main_kind ? d_lm : d_lb;
        a.lut_off = main_kind ? d_offm : d_offb;
        for (int s = 0; s <= shares; ++s) a.first[s] = first[s];
        a.shares = shares;
        a.counts = main_kind ? d_cm : d_cb;
        a.offs = main_kind ? d_om : d_ob;
        a.totals = d_tot + (main_kind ? 0 : kMaxShares);
        a.out_idx = main_kind ? d_im : d_ib;
        a.out_res = main_kind ? d_rm : d_rb;
        return a;
    };
    const XchgLaunch am = launch_of(true), ab = launch_of(false);
    // The main and blocked words' launches are independent (own tiles, counts,
    // totals and buckets): the blocked ones run on a second stream beside them
    // where the group sweeps' streams exist.
    hipStream_t xs = c->gfork ? c->gstream[0] : c->stream;
    OFP_HIP(c, hipEventRecord(c->ev0, c->stream), "event");
    if (xs != c->stream) {
        OFP_HIP(c, hipEventRecord(c->gfork, c->stream), "exchange fork");
        OFP_HIP(c, hipStreamWaitEvent(xs, c->gfork, 0), "exchange fork");
    }
    OFP_HIP(c, launch_exchange_count(ab, xs), "exchange count");
    OFP_HIP(c, launch_exchange_scan(ab, xs), "exchange scan");
    OFP_HIP(c, launch_exchange_write(ab, xs), "exchange write");
    OFP_HIP(c, launch_exchange_count(am, c->stream), "exchange count");
    OFP_HIP(c, launch_exchange_scan(am, c->stream), "exchange scan");
    OFP_HIP(c, launch_exchange_write(am, c->stream), "exchange write");
    if (xs != c->stream) {
        OFP_HIP(c, hipEventRecord(c->gjoin[0], xs), "exchange join");
        OFP_HIP(c, hipStreamWaitEvent(c->stream, c->gjoin[0], 0), "exchange join");
    }
    OFP_HIP(c, hipEventRecord(c->ev1, c->stream), "event");
    OFP_HIP(c, hipEventSynchronize(c->ev1), "exchange sync");
    float ms = 0;
    OFP_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1), "event time");
    uint64_t tot[2 * kMaxShares] = {};
    OFP_HIP(c, hipMemcpy(tot, d_tot, sizeof tot, hipMemcpyDeviceToHost), "read bucket counts");
    *info = ofp_exchange_info{};
    info->main_count = M;
    info->blocked_count = B;
    info->shares = shares;
    uint64_t sm = 0, sb = 0;
    for (int s = 0; s <= shares; ++s) info->share_first[s] = first[s];
    for (int s = 0; s < shares; ++s) {
        info->main_records[s] = tot[s];
        info->blocked_records[s] = tot[kMaxShares + s];
        sm += tot[s];
        sb += tot[kMaxShares + s];
    }
    if (sm != M || sb != B) return fail(c, OFP_E_RUNTIME, "exchange bucket counts do not add up to the window");
    info->staging_bytes = b_im + b_ib + b_rm + b_rb;
    info->moved_bytes = (M + B) * (uint64_t)Kp * 4 + (M + B) * (8 + (uint64_t)Kp * 4);
    info->kernel_s = ms * 1e-3;
    if (main_index && M) OFP_HIP(c, hipMemcpy(main_index, d_im, M * 8, hipMemcpyDeviceToHost), "download records");
    if (blocked_index && B)
        OFP_HIP(c, hipMemcpy(blocked_index, d_ib, B * 8, hipMemcpyDeviceToHost), "download records");
    if ((main_res && M) || (blocked_res && B)) {
        std::vector<uint32_t> t(std::max(M, B) * (size_t)Kp);
        if (main_res && M) {
            OFP_HIP(c, hipMemcpy(t.data(), d_rm, M * Kp * 4, hipMemcpyDeviceToHost), "download records");
            for (uint64_t i = 0; i < M; ++i)
                for (int j = 0; j < k; ++j) main_res[i * k + j] = t[i * Kp + j];
        }
        if (blocked_res && B) {
            OFP_HIP(c, hipMemcpy(t.data(), d_rb, B * Kp * 4, hipMemcpyDeviceToHost), "download records");
            for (uint64_t i = 0; i < B; ++i)
                for (int j = 0; j < k; ++j) blocked_res[i * k + j] = t[i * Kp + j];
        }
    }
    return OFP_OK;
}

// ---- jobs: independent device-resident solves that overlap on the GPU ----

ofp_status ofp_job_create(ofp_ctx* c, int n, const uint32_t* moduli, int k, ofp_job** out) {
    if (!c || !moduli || !out) return OFP_E_ARG;
    *out = nullptr;
    ofp_job* j = new ofp_job();
    j->ctx = c;
    auto cleanup = [&](ofp_status st) {
        free_sweep(j->sw);
        if (j->sw.stream) (void)hipStreamDestroy(j->sw.stream);
        if (j->done) (void)hipEventDestroy(j->done);
        delete j;
        return st;
    };
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return cleanup(hip_fail(c, e, "hipSetDevice"));
    e = hipStreamCreateWithFlags(&j->sw.stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(hip_fail(c, e, "job stream"));
    e = hipEventCreateWithFlags(&j->done, hipEventDisableTiming);
    if (e != hipSuccess) return cleanup(hip_fail(c, e, "job event"));
    j->sw.async_alloc = true;
    ofp_status st = ensure_state(c, j->sw, n, moduli, k);
    if (st) return cleanup(st);
    st = init_state(c, j->sw);
    if (st) return cleanup(st);
    e = hipEventRecord(j->done, j->sw.stream);
    if (e != hipSuccess) return cleanup(hip_fail(c, e, "job event record"));
    ++c->jobs;
    *out = j;
    return OFP_OK;
}

ofp_status ofp_job_enqueue(ofp_job* j, int rows) {
    if (!j || rows < 0) return OFP_E_ARG;
    ofp_ctx* c = j->ctx;
    if (j->rows_enqueued + rows > j->sw.W) return fail(c, OFP_E_ARG, "job has only W rows");
    ofp_status st = run_rows(c, j->sw, rows, nullptr);
    if (st) return st;
    j->rows_enqueued += rows;
    OFP_HIP(c, hipEventRecord(j->done, j->sw.stream), "job event record");
    return OFP_OK;
}

ofp_status ofp_job_info_get(ofp_job* j, ofp_job_info* info) {
    if (!j || !info) return OFP_E_ARG;
    hipError_t q = hipEventQuery(j->done);
    if (q != hipSuccess && q != hipErrorNotReady) return hip_fail(j->ctx, q, "job query");
    info->n = j->sw.n;
    info->k = j->sw.K;
    info->rows_total = j->sw.W;
    info->rows_enqueued = j->rows_enqueued;
    info->idle = (q == hipSuccess) ? 1 : 0;
    info->state_bytes = j->sw.main_cap + j->sw.blk_cap;
    info->transitions_per_row = (uint64_t)(j->sw.W - 1) * (j->sw.nmain + j->sw.nblk) * (uint64_t)j->sw.K;
    return OFP_OK;
}

ofp_status ofp_job_result(ofp_job* j, uint32_t* residues_out) {
    if (!j || !residues_out) return OFP_E_ARG;
    if (j->rows_enqueued != j->sw.W) return fail(j->ctx, OFP_E_ARG, "job has not run all W rows");
    return read_main(j->ctx, j->sw, 0, residues_out, j->sw.K);
}

void ofp_job_destroy(ofp_job* j) {
    if (!j) return;
    (void)hipSetDevice(j->ctx->device);
    if (j->sw.stream) (void)hipStreamSynchronize(j->sw.stream);
    free_sweep(j->sw);
    if (j->sw.stream) (void)hipStreamDestroy(j->sw.stream);
    if (j->done) (void)hipEventDestroy(j->done);
    --j->ctx->jobs;
    delete j;
}

void ofp_destroy(ofp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    free_sweep(c->st);
    if (c->dp32) (void)hipFree(c->dp32);
    if (c->dp64) (void)hipFree(c->dp64);
    if (c->dpc32) (void)hipFree(c->dpc32);
    if (c->dpc64) (void)hipFree(c->dpc64);
    if (c->ct32) (void)hipFree(c->ct32);
    if (c->ct64) (void)hipFree(c->ct64);
    for (int i = 0; i < 3; ++i) {
        if (c->gstream[i]) (void)hipStreamDestroy(c->gstream[i]);
        if (c->gjoin[i]) (void)hipEventDestroy(c->gjoin[i]);
    }
    if (c->gfork) (void)hipEventDestroy(c->gfork);
    if (c->win) (void)hipFree(c->win);
    if (c->stage) (void)hipFree(c->stage);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
