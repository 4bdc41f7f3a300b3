// ofp_cli.cpp -- drop-in CLI for the reference's GPU solver binary.
//
// argv and stdout follow the reference process contract:
//   <binary> n target_mib max_window ngpu mod1 [mod2 ...]
//   (src/cuda/b300/oneesan_cuda_gridfp_b300_hbm32_factorized_reverse2_row6crt20_batch.cu:1690-1701;
//    identical in ..._fullmate_dropN_batch.cu:467-478; pinned by tests/test_solver_contract.py:21-31)
// with the reference defaults n=16, target_mib=16384, max_window=14, ngpu=all
// visible, modulus 4294967291.  One stdout line per modulus, in argument
// order, beginning with backend= and carrying n=, residue=, modulus=,
// wall_s= exactly once each and in that order (row6crt20.cu:1841).
//
// Differences, by design: the moduli are computed together in interleaved
// sweeps -- all in one when their state fits HBM, else in sweeps of the largest
// count that fits, one after another (the library's ofp_solve; the environment
// variable OFP_BATCH_MODULI caps the count per sweep) -- so each line reports
// wall_s = the whole run's wall time / k (its share, so that summing wall_s over a
// run gives the solver's real time, as the exact runner's manifest check expects)
// and batch_seconds = the whole run; target_mib and max_window are
// validated and echoed but do not change the sweep (there is no scratch
// window: the state is updated in place); one GPU is used.
// Exit codes: 1 bad arguments / runtime error, 2 no GPU, 5 state does not fit
// HBM (reference: row6crt20.cu:1698-1790).
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/ofp.h"

static bool parse_int(const char* s, long lo, long hi, long* out) {
    errno = 0;
    char* end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (errno || !end || *end || end == s || v < lo || v > hi) return false;
    *out = v;
    return true;
}

int main(int argc, char** argv) {
    long n = 16, target_mib = 16384, max_window = 14, requested = 0;
    if (argc > 1 && !parse_int(argv[1], 1, OFP_MAX_N, &n)) {
        std::cerr << "n must be an integer in 1.." << OFP_MAX_N << "\n";
        return 1;
    }
    if (argc > 2 && !parse_int(argv[2], 0, 1L << 30, &target_mib)) {
        std::cerr << "target_mib must be a non-negative integer\n";
        return 1;
    }
    if (argc > 3 && !parse_int(argv[3], 1, 64, &max_window)) {
        std::cerr << "max_window must be an integer in 1..64\n";
        return 1;
    }
    if (argc > 4 && !parse_int(argv[4], 0, 64, &requested)) {
        std::cerr << "ngpu must be an integer in 0..64\n";
        return 1;
    }
    std::vector<uint32_t> mods;
    for (int i = 5; i < argc; ++i) {
        errno = 0;
        char* end = nullptr;
        unsigned long long raw = std::strtoull(argv[i], &end, 10);
        if (errno || !end || *end || argv[i][0] == '-' || raw < 2 || raw > 0xffffffffULL) {
            std::cerr << "HBM32 modulus must be in [2, 4294967295], got " << argv[i] << "\n";
            return 1;
        }
        mods.push_back((uint32_t)raw);
    }
    if (mods.empty()) mods.push_back(4294967291u);

    ofp_config cfg{};
    cfg.device = 0;
    cfg.use_graphs = 1;
    if (const char* e = std::getenv("GRIDFP_VRAM_RESERVE_MIB")) {
        long v;
        if (!parse_int(e, 0, 1L << 24, &v)) {
            std::cerr << "GRIDFP_VRAM_RESERVE_MIB must be a non-negative integer\n";
            return 1;
        }
        cfg.reserve_bytes = (size_t)v << 20;
    }
    if (const char* e = std::getenv("GRIDFP_FRONTIER_THREADS")) {
        long v;
        if (!parse_int(e, 64, 256, &v) || (v & 63)) {
            std::cerr << "GRIDFP_FRONTIER_THREADS must be 64, 128, 192 or 256\n";
            return 1;
        }
        cfg.threads = (int)v;
    }
    ofp_ctx* ctx = nullptr;
    ofp_status st = ofp_create(&cfg, &ctx);
    if (st == OFP_E_DEVICE) {
        std::cerr << "need 1 GPU\n";
        return 2;
    }
    if (st != OFP_OK) {
        std::cerr << "ofp_create failed (" << st << ")\n";
        return 1;
    }
    if (const char* e = std::getenv("OFP_BATCH_MODULI")) {
        long v;
        if (!parse_int(e, 0, 1024, &v) || ofp_set_max_batch(ctx, (int)v) != OFP_OK) {
            std::cerr << "OFP_BATCH_MODULI must be an integer in 0..1024\n";
            ofp_destroy(ctx);
            return 1;
        }
    }
    ofp_plan_info plan{};
    if (ofp_plan(ctx, (int)n, (int)mods.size(), &plan) != OFP_OK) {
        std::cerr << ofp_last_error(ctx) << "\n";
        ofp_destroy(ctx);
        return 1;
    }
    std::cerr << "ofp state: main_states=" << plan.main_states << " blocked_states=" << plan.blocked_states
              << " moduli=" << mods.size() << " state_gib=" << double(plan.state_bytes) / (1ull << 30)
              << " free_gib=" << double(plan.free_bytes) / (1ull << 30) << "\n";
    if (!plan.fits) {
        ofp_plan_info one{};
        if (ofp_plan(ctx, (int)n, 1, &one) != OFP_OK || !one.fits) {
            std::cerr << "insufficient HBM for the state of one modulus\n";
            ofp_destroy(ctx);
            return 5;
        }
        std::cerr << "the state of all moduli does not fit: they run in several sweeps\n";
    }
    std::vector<uint32_t> res(mods.size());
    ofp_stats stats{};
    st = ofp_solve(ctx, (int)n, mods.data(), (int)mods.size(), res.data(), &stats);
    if (st != OFP_OK) {
        std::cerr << "solve failed: " << ofp_last_error(ctx) << "\n";
        ofp_destroy(ctx);
        return st == OFP_E_MEMORY ? 5 : 1;
    }
    const size_t sweeps = stats.rows > 0 ? (size_t)stats.rows / (size_t)(n + 1) : 1;
    for (size_t j = 0; j < mods.size(); ++j) {
        std::cout << "backend=ofp-mi355x-hbm32-interleaved n=" << n << " residue=" << res[j]
                  << " modulus=" << mods[j] << " residue_index=" << j << " residues_total=" << mods.size()
                  << " gpus=1 peers=0 main_states=" << plan.main_states
                  << " blocked_states=" << plan.blocked_states << " scratch_target_mib=" << target_mib
                  << " max_window=" << max_window << " requested_gpus=" << requested
                  << " batch_moduli=" << (mods.size() + sweeps - 1) / sweeps << " sweeps=" << sweeps
                  << " transitions=" << stats.transitions
                  << " kernel_s=" << stats.kernel_s << " launches=" << stats.launches
                  << " batch_seconds=" << stats.wall_s
                  << " wall_s=" << stats.wall_s / double(mods.size()) << std::endl;
    }
    ofp_destroy(ctx);
    return 0;
}
