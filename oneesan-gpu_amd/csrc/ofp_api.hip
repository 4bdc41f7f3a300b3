// ofp_api.hip -- host runtime behind include/ofp.h.
//
// One context = one GPU, one non-blocking stream (default priority), the
// height table on the device, and one device-resident frontier state
// (main + blocked words, k moduli interleaved in 64-state tiles).
// A solve initialises the state (reference: fullmate_dropN_batch.cu:508-511),
// runs W rows of cells p = W-1..1 (each row captured once in a hipGraph and
// replayed), and reads main[rank(R)] = main[0] (fullmate_dropN_batch.cu:514).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ofp.h"
#include "ofp_internal.h"

using namespace ofp;

struct ofp_ctx {
    int device = 0;
    int cus = 256;
    int threads = 256;
    int blocks_per_cu = 32;
    int use_graphs = 1;
    int layout = 1;  // 0: 64-state tiles, 1: state-major with even stride
    size_t reserve = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* dp32 = nullptr;  // packed uint32 height table (n <= 22)
    void* dp64 = nullptr;  // packed uint64 height table
    std::string err;
    // device-resident state
    uint32_t* main = nullptr;
    uint32_t* blk = nullptr;
    size_t main_cap = 0, blk_cap = 0;  // bytes allocated
    uint32_t* mods_dev = nullptr;
    int n = 0, W = 0, K = 0, Kp = 0;  // k moduli, interleave stride Kp
    int state_layout = -1;           // layout the current buffers were written in
    std::vector<uint32_t> mods;
    uint64_t nmain = 0, nblk = 0;
    // row graph cache
    hipGraphExec_t row_exec = nullptr;
    hipGraph_t row_graph = nullptr;
    std::string row_key;
};

namespace {

const char* kVersion = "ofp-mi355x 0.1 (per-cell in-place orbit sweep, gfx950)";

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

void drop_graph(ofp_ctx* c) {
    if (c->row_exec) (void)hipGraphExecDestroy(c->row_exec);
    if (c->row_graph) (void)hipGraphDestroy(c->row_graph);
    c->row_exec = nullptr;
    c->row_graph = nullptr;
    c->row_key.clear();
}

int grid_blocks(const ofp_ctx* c, uint64_t nblk) {
    uint64_t chunks = (nblk + c->threads - 1) / c->threads;
    uint64_t cap = (uint64_t)c->cus * c->blocks_per_cu;
    uint64_t b = std::min<uint64_t>(chunks, cap);
    b = (b + 7) / 8 * 8;
    return (int)std::max<uint64_t>(b, 8);
}

// Ensure device buffers for (n, moduli) and record them as the current state.
ofp_status ensure_state(ofp_ctx* c, int n, const uint32_t* moduli, int k) {
    uint64_t nm, nb;
    if (!counts(n, &nm, &nb)) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    if (k < 1 || k > 1024) return fail(c, OFP_E_ARG, "number of moduli must be in 1..1024");
    for (int j = 0; j < k; ++j)
        if (moduli[j] < 2) return fail(c, OFP_E_ARG, "modulus must be in [2, 4294967295]");
    const int Kp = interleave_stride(c->layout, k);
    size_t mb = padded_states(nm) * (size_t)Kp * sizeof(uint32_t);
    size_t bb = padded_states(nb ? nb : 1) * (size_t)Kp * sizeof(uint32_t);
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    if (mb > c->main_cap || bb > c->blk_cap) {
        drop_graph(c);
        if (c->main) (void)hipFree(c->main);
        if (c->blk) (void)hipFree(c->blk);
        c->main = nullptr;
        c->blk = nullptr;
        c->main_cap = c->blk_cap = 0;
        size_t fr = 0, tot = 0;
        OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
        if (mb + bb + c->reserve > fr)
            return fail(c, OFP_E_MEMORY, "insufficient HBM: need " + std::to_string(mb + bb) +
                                             " bytes, free " + std::to_string(fr));
        OFP_HIP(c, hipMalloc(&c->main, mb), "hipMalloc main state");
        OFP_HIP(c, hipMalloc(&c->blk, bb), "hipMalloc blocked state");
        c->main_cap = mb;
        c->blk_cap = bb;
    }
    if (c->mods_dev == nullptr || (int)c->mods.size() < k) {
        if (c->mods_dev) (void)hipFree(c->mods_dev);
        c->mods_dev = nullptr;
        OFP_HIP(c, hipMalloc(&c->mods_dev, std::max(k, 16) * sizeof(uint32_t)), "hipMalloc moduli");
    }
    c->mods.assign(moduli, moduli + k);
    OFP_HIP(c, hipMemcpy(c->mods_dev, moduli, k * sizeof(uint32_t), hipMemcpyHostToDevice),
            "copy moduli");
    c->n = n;
    c->W = n + 1;
    c->K = k;
    c->Kp = Kp;
    c->state_layout = c->layout;
    c->nmain = nm;
    c->nblk = nb;
    return OFP_OK;
}

// Enqueue cells p_hi..p_lo on the context stream.
ofp_status enqueue_cells(ofp_ctx* c, int p_hi, int p_lo, bool first_reads_blk, int* launches) {
    const bool wide = c->W > 23;
    const int blocks = grid_blocks(c, c->nblk);
    for (int p = p_hi; p >= p_lo; --p) {
        for (int off = 0; off < c->K; off += kMaxKC) {
            CellLaunch a{};
            a.main = c->main;
            a.blk = c->blk;
            a.dp = wide ? c->dp64 : c->dp32;
            a.nblk = c->nblk;
            a.first = 0;
            a.layout = c->layout;
            a.W = c->W;
            a.p = p;
            a.read_blk = (p == p_hi) ? (first_reads_blk ? 1 : 0) : 1;
            a.K = c->Kp;
            a.off = off;
            a.kc = std::min(kMaxKC, c->K - off);
            for (int j = 0; j < a.kc; ++j) a.mod[j] = c->mods[off + j];
            hipError_t e = launch_cell(a, wide, c->threads, blocks, c->stream);
            if (e != hipSuccess) return hip_fail(c, e, "cell kernel launch");
            if (launches) ++*launches;
        }
    }
    return OFP_OK;
}

// One full row (cells W-1..1, blocked input of cell W-1 is zero), via a cached graph.
ofp_status run_rows(ofp_ctx* c, int rows, int* launches) {
    if (!c->use_graphs) {
        for (int r = 0; r < rows; ++r) {
            ofp_status s = enqueue_cells(c, c->W - 1, 1, false, launches);
            if (s) return s;
        }
        return OFP_OK;
    }
    char key[256];
    snprintf(key, sizeof key, "%p/%p/%d/%d/%d/%d/%d", (void*)c->main, (void*)c->blk, c->n, c->K,
             c->threads, c->blocks_per_cu, c->layout);
    std::string k = key;
    for (uint32_t m : c->mods) k += "/" + std::to_string(m);
    if (k != c->row_key || !c->row_exec) {
        drop_graph(c);
        OFP_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "begin capture");
        int dummy = 0;
        ofp_status s = enqueue_cells(c, c->W - 1, 1, false, &dummy);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (s) { if (g) (void)hipGraphDestroy(g); return s; }
        if (e != hipSuccess) return hip_fail(c, e, "end capture");
        c->row_graph = g;
        OFP_HIP(c, hipGraphInstantiate(&c->row_exec, g, nullptr, nullptr, 0), "graph instantiate");
        c->row_key = k;
    }
    for (int r = 0; r < rows; ++r) {
        OFP_HIP(c, hipGraphLaunch(c->row_exec, c->stream), "graph launch");
        if (launches) *launches += (c->W - 1) * ((c->K + kMaxKC - 1) / kMaxKC);
    }
    return OFP_OK;
}

ofp_status init_state(ofp_ctx* c) {
    OFP_HIP(c, hipMemsetAsync(c->main, 0, padded_states(c->nmain) * (size_t)c->Kp * 4, c->stream),
            "zero main state");
    // start word: R at position W-1, rank = H_DP[W-1][1] (fullmate_dropN_batch.cu:508)
    uint64_t start = dp_at(c->W - 1, 1);
    hipError_t e = launch_set_one(c->main, start, c->Kp, c->K, c->layout, c->mods_dev, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "init kernel");
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
    if (c->threads <= 0 || c->threads > 256 || (c->threads & 63) || c->blocks_per_cu <= 0) {
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
        size_t r = std::min<size_t>(8ull << 30, std::max<size_t>(256ull << 20, tot / 32));
        c->reserve = r;
    }
    if (bad(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking))) return OFP_E_RUNTIME;
    if (bad(hipEventCreate(&c->ev0)) || bad(hipEventCreate(&c->ev1))) return OFP_E_RUNTIME;
    std::vector<unsigned char> q32(64 * 1024), q64(64 * 1024);
    size_t b32 = pack_height_table(false, q32.data(), q32.size());
    size_t b64 = pack_height_table(true, q64.data(), q64.size());
    if (!b32 || !b64) {
        delete c;
        return OFP_E_RUNTIME;
    }
    if (bad(hipMalloc(&c->dp32, b32)) || bad(hipMalloc(&c->dp64, b64))) return OFP_E_RUNTIME;
    if (bad(hipMemcpy(c->dp32, q32.data(), b32, hipMemcpyHostToDevice)) ||
        bad(hipMemcpy(c->dp64, q64.data(), b64, hipMemcpyHostToDevice)))
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
    if (threads != c->threads || blocks_per_cu != c->blocks_per_cu) drop_graph(c);
    c->threads = threads;
    c->blocks_per_cu = blocks_per_cu;
    return OFP_OK;
}

ofp_status ofp_set_layout(ofp_ctx* c, int layout) {
    if (!c) return OFP_E_ARG;
    if (layout != 0 && layout != 1) return fail(c, OFP_E_ARG, "layout must be 0 (64-state tiles) or 1 (state-major)");
    if (layout != c->layout) {
        drop_graph(c);
        c->layout = layout;
        c->state_layout = -1;  // current contents no longer readable; prepare again
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
    out->state_bytes = (padded_states(nm) + padded_states(nb ? nb : 1)) *
                       (uint64_t)interleave_stride(c->layout, k) * 4;
    const uint64_t W = n + 1;
    out->transitions = W * (W - 1) * (nm + nb) * (uint64_t)k;
    size_t fr = 0, tot = 0;
    OFP_HIP(c, hipSetDevice(c->device), "hipSetDevice");
    OFP_HIP(c, hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
    // memory already held by this context's state can be reused
    uint64_t reusable = c->main_cap + c->blk_cap;
    out->free_bytes = fr + reusable;
    out->fits = out->state_bytes + c->reserve <= out->free_bytes;
    return OFP_OK;
}

ofp_status ofp_prepare(ofp_ctx* c, int n, const uint32_t* moduli, int k) {
    if (!c || !moduli) return OFP_E_ARG;
    ofp_status s = ensure_state(c, n, moduli, k);
    if (s) return s;
    s = init_state(c);
    if (s) return s;
    OFP_HIP(c, hipStreamSynchronize(c->stream), "prepare sync");
    return OFP_OK;
}

ofp_status ofp_row_sweep(ofp_ctx* c, int rows, ofp_stats* st) {
    if (!c || rows < 0) return OFP_E_ARG;
    if (!c->main || c->state_layout != c->layout) return fail(c, OFP_E_ARG, "no prepared state");
    int launches = 0;
    auto t0 = std::chrono::steady_clock::now();
    OFP_HIP(c, hipEventRecord(c->ev0, c->stream), "event record");
    ofp_status s = run_rows(c, rows, &launches);
    if (s) return s;
    OFP_HIP(c, hipEventRecord(c->ev1, c->stream), "event record");
    OFP_HIP(c, hipEventSynchronize(c->ev1), "row sweep sync");
    float ms = 0;
    OFP_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1), "event time");
    if (st) {
        st->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        st->kernel_s = ms * 1e-3;
        st->main_states = c->nmain;
        st->blocked_states = c->nblk;
        st->rows = rows;
        st->cells_per_row = c->W - 1;
        st->launches = launches;
        st->transitions = (uint64_t)rows * (c->W - 1) * (c->nmain + c->nblk) * (uint64_t)c->K;
    }
    return OFP_OK;
}

ofp_status ofp_read_main(ofp_ctx* c, uint64_t index, uint32_t* out_k) {
    if (!c || !out_k || !c->main) return OFP_E_ARG;
    if (index >= c->nmain) return fail(c, OFP_E_ARG, "index out of range");
    OFP_HIP(c, hipStreamSynchronize(c->stream), "sync");
    for (int j = 0; j < c->K; ++j)
        OFP_HIP(c, hipMemcpy(out_k + j, c->main + state_elem(c->state_layout, index, c->Kp, j), 4,
                             hipMemcpyDeviceToHost),
                "read residue");
    return OFP_OK;
}

ofp_status ofp_solve(ofp_ctx* c, int n, const uint32_t* moduli, int k, uint32_t* residues_out,
                     ofp_stats* st) {
    if (!c || !moduli || !residues_out) return OFP_E_ARG;
    auto t0 = std::chrono::steady_clock::now();
    ofp_status s = ensure_state(c, n, moduli, k);
    if (s) return s;
    s = init_state(c);
    if (s) return s;
    ofp_stats rs{};
    s = ofp_row_sweep(c, c->W, &rs);
    if (s) return s;
    s = ofp_read_main(c, 0, residues_out);  // rank_full(R) = 0 (fullmate_dropN_batch.cu:508,514)
    if (s) return s;
    if (st) {
        *st = rs;
        st->wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    return OFP_OK;
}

ofp_status ofp_apply_cells(ofp_ctx* c, int n, const uint32_t* moduli, int k, uint32_t* main_io,
                           uint32_t* blocked_io, int p_hi, int p_lo, int blocked_valid) {
    if (!c || !moduli || !main_io || !blocked_io) return OFP_E_ARG;
    if (n < 1 || n > OFP_MAX_N) return fail(c, OFP_E_RANGE, "n must be in 1..27");
    const int W = n + 1;
    if (p_hi > W - 1 || p_lo < 1 || p_hi < p_lo) return fail(c, OFP_E_ARG, "need W-1 >= p_hi >= p_lo >= 1");
    ofp_status s = ensure_state(c, n, moduli, k);
    if (s) return s;
    // host state-major -> device tiles
    const int Kp = c->Kp, lay = c->layout;
    std::vector<uint32_t> tm(padded_states(c->nmain) * (size_t)Kp, 0),
        tb(padded_states(c->nblk ? c->nblk : 1) * (size_t)Kp, 0);
    for (uint64_t i = 0; i < c->nmain; ++i)
        for (int j = 0; j < k; ++j) tm[state_elem(lay, i, Kp, j)] = main_io[i * k + j];
    for (uint64_t i = 0; i < c->nblk; ++i)
        for (int j = 0; j < k; ++j) tb[state_elem(lay, i, Kp, j)] = blocked_io[i * k + j];
    OFP_HIP(c, hipMemcpy(c->main, tm.data(), tm.size() * 4, hipMemcpyHostToDevice), "upload main");
    OFP_HIP(c, hipMemcpy(c->blk, tb.data(), tb.size() * 4, hipMemcpyHostToDevice), "upload blocked");
    s = enqueue_cells(c, p_hi, p_lo, blocked_valid != 0, nullptr);
    if (s) return s;
    OFP_HIP(c, hipStreamSynchronize(c->stream), "apply sync");
    OFP_HIP(c, hipMemcpy(tm.data(), c->main, tm.size() * 4, hipMemcpyDeviceToHost), "download main");
    OFP_HIP(c, hipMemcpy(tb.data(), c->blk, tb.size() * 4, hipMemcpyDeviceToHost), "download blocked");
    for (uint64_t i = 0; i < c->nmain; ++i)
        for (int j = 0; j < k; ++j) main_io[i * k + j] = tm[state_elem(lay, i, Kp, j)];
    // after cell 1 nothing is blocked; the device buffer is simply not written
    for (uint64_t i = 0; i < c->nblk; ++i)
        for (int j = 0; j < k; ++j) blocked_io[i * k + j] = (p_lo == 1) ? 0u : tb[state_elem(lay, i, Kp, j)];
    return OFP_OK;
}

void ofp_destroy(ofp_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    drop_graph(c);
    if (c->main) (void)hipFree(c->main);
    if (c->blk) (void)hipFree(c->blk);
    if (c->mods_dev) (void)hipFree(c->mods_dev);
    if (c->dp32) (void)hipFree(c->dp32);
    if (c->dp64) (void)hipFree(c->dp64);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

}  // extern "C"
