/*
 * ofp.h -- C ABI of the MI355X frontier transfer-matrix solver (libofp.so).
 *
 * Counts simple corner-to-corner paths on an (n+1)x(n+1) vertex grid
 * (OEIS A007764) modulo one or more uint32 moduli, with the whole
 * frontier state (main + blocked words, all moduli interleaved per state)
 * resident in one GPU's HBM.
 *
 * The reference has no library boundary: its GPU path is a process whose
 * argv is `n target_mib max_window ngpu mod...` and whose stdout carries one
 * `backend=... n= residue= modulus= ... wall_s=` line per modulus
 * (src/cuda/b300/oneesan_cuda_gridfp_b300_hbm32_factorized_reverse2_row6crt20_batch.cu:1690-1701,1841;
 *  src/cuda/b300/oneesan_cuda_gridfp_b300_hbm32_fullmate_dropN_batch.cu:467-514).
 * Each entry point below names the part of that process it replaces.  The
 * CLI in oneesan-gpu_amd/csrc/ofp_cli.cpp reproduces the process contract on
 * top of this ABI.
 *
 * Conventions: every function returns an ofp_status (0 = OFP_OK, negative on
 * error) and never exits the process; ofp_last_error() gives the message.
 * The caller owns host buffers; the library owns device memory.  A context
 * is single-threaded (one host thread at a time).
 */
#ifndef OFP_H
#define OFP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t ofp_status;
#define OFP_OK 0
#define OFP_E_ARG (-1)      /* bad argument (reference exit code 1) */
#define OFP_E_DEVICE (-2)   /* no usable GPU (reference exit code 2) */
#define OFP_E_MEMORY (-5)   /* state does not fit HBM (reference exit code 5) */
#define OFP_E_RUNTIME (-6)  /* HIP runtime error (reference ck() -> exit 1) */
#define OFP_E_RANGE (-7)    /* n outside the supported range */

#define OFP_MAX_N 27        /* W = n+1 <= 28, as the reference's MAXW */

typedef struct ofp_ctx ofp_ctx;

typedef struct ofp_config {
    int device;             /* HIP device ordinal */
    int threads;            /* threads per workgroup, 0 = default (256) */
    int blocks_per_cu;      /* workgroups per CU for the cell kernel, 0 = default */
    size_t reserve_bytes;   /* HBM to leave free (reference GRIDFP_VRAM_RESERVE_MIB), 0 = default */
    int use_graphs;         /* capture each row sweep in a hipGraph (default 1) */
    int layout;             /* state layout: 0 = default, 1 = 64-state tiles, 2 = state-major */
} ofp_config;

typedef struct ofp_plan_info {
    uint64_t main_states;     /* H_DP[W][1] = M_{n+2} - M_{n+1}  (row6crt20.cu:1748) */
    uint64_t blocked_states;  /* H_DP[W-1][1] */
    uint64_t state_bytes;     /* device bytes of the interleaved state for k moduli */
    uint64_t transitions;     /* W*(W-1)*(main+blocked)*k: frontier transitions per solve */
    uint64_t free_bytes;      /* free HBM on the device right now */
    int fits;                 /* state_bytes + reserve <= free_bytes */
} ofp_plan_info;

typedef struct ofp_stats {
    double wall_s;            /* whole solve, host clock */
    double kernel_s;          /* sum of cell-kernel time (HIP events), 0 if not timed */
    uint64_t transitions;     /* frontier transitions performed */
    uint64_t main_states, blocked_states;
    int rows, cells_per_row, launches;
    /* kernel time split: runs of rows whose cells all sweep the full state
     * (full_*) and rows with height-capped cells (capped_*) */
    double full_kernel_s, capped_kernel_s;
    int full_launches, capped_launches;
} ofp_stats;

/* Create a context on cfg->device (NULL = defaults).  Replaces the
 * reference's device discovery and DeviceCtx::init
 * (row6crt20.cu:1716-1750; fullmate_dropN_batch.cu:483-497). */
ofp_status ofp_create(const ofp_config* cfg, ofp_ctx** out);

/* Predict state size and work for (n, k).  Replaces the reference's HBM
 * budget check (row6crt20.cu:1751-1755; fullmate_dropN_batch.cu:495-501);
 * the serving loop admits requests with it. */
ofp_status ofp_plan(ofp_ctx* ctx, int n, int k, ofp_plan_info* out);

/* Full solve: residues_out[j] = A007764(n) mod moduli[j] for j < k, all k
 * moduli advanced together as one interleaved residue vector per state.
 * Replaces the reference's residue loop (row6crt20.cu:1784-1841;
 * fullmate_dropN_batch.cu:509-515), which ran the moduli one after another.
 * moduli[j] must be in [2, 2^32-1] (row6crt20.cu:1697-1698). */
ofp_status ofp_solve(ofp_ctx* ctx, int n, const uint32_t* moduli, int k,
                     uint32_t* residues_out, ofp_stats* stats);

/* ofp_solve for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation): the
 * same result, with the state held in a static device array inside the library
 * (64 MiB per device the library runs on, kept until the process ends) and no
 * hipMalloc, hipFree or graph creation on the way.  The state of (n, k) must fit
 * the array (n <= 15 with up to 8 moduli, n = 16 with up to 2); larger requests
 * fail with OFP_E_MEMORY; 1..16 moduli.  Each call re-initialises the array, and
 * calls that overlap on one device (several contexts or threads) race on it.
 * No counterpart in the reference, which allocates its state per process. */
ofp_status ofp_solve_static(ofp_ctx* ctx, int n, const uint32_t* moduli, int k,
                            uint32_t* residues_out);

/* Moduli per sweep of ofp_solve.  A solve whose k-moduli state does not fit
 * HBM (or exceeds max_moduli, when that is > 0) runs its moduli in several
 * sweeps one after another, each of the largest count that fits, split as
 * ofp_moduli_groups() says; the residues are the same.  0 (default) = as many
 * as fit.  Reference analogue: its residue loop runs one modulus per sweep
 * (row6crt20.cu:1784-1841). */
ofp_status ofp_set_max_batch(ofp_ctx* ctx, int max_moduli);

/* The sweeps ofp_solve makes of k moduli with at most max_per_sweep each:
 * fewest sweeps, near-equal sizes, even sizes where that adds no sweep.
 * Writes up to `capacity` sizes and returns the number of sweeps (-1 on bad
 * arguments).  Host-only, no GPU. */
int ofp_moduli_groups(int k, int max_per_sweep, int* sizes, int capacity);

/* Apply cells p_hi..p_lo (W-1 >= p_hi >= p_lo >= 1) of one row to caller
 * state, on the device, in the product layout.  main_io holds
 * main_states*k residues and blocked_io blocked_states*k residues, both
 * state-major (index*k + j) on the host; they are updated in place.  If
 * blocked_valid is 0, the blocked input of the first cell is taken as zero
 * (as at the start of a row).  This is the unit the reference's own GPU
 * tests drive (tests/factor_transfer_batch_test.cu:44-96: one window of
 * cells on seeded random vectors, compared with a CPU forward recurrence). */
ofp_status ofp_apply_cells(ofp_ctx* ctx, int n, const uint32_t* moduli, int k,
                           uint32_t* main_io, uint32_t* blocked_io,
                           int p_hi, int p_lo, int blocked_valid);

/* One timed row sweep (cells W-1..1) over the device-resident state of the
 * last ofp_prepare() call; used by the benchmark and the autotuner.  The
 * state is not re-initialised, so repeated calls advance it row by row. */
ofp_status ofp_prepare(ofp_ctx* ctx, int n, const uint32_t* moduli, int k);
ofp_status ofp_row_sweep(ofp_ctx* ctx, int rows, ofp_stats* stats);
/* Read the residues of main state `index` from the prepared state: one per
 * modulus of the last ofp_prepare() / ofp_solve(), in order.  `capacity` is
 * the number of uint32 slots at out; fewer than the prepared moduli count is
 * refused with OFP_E_ARG (nothing is written).  The answer word is index 0
 * (rank_full(R), fullmate_dropN_batch.cu:514). */
ofp_status ofp_read_main(ofp_ctx* ctx, uint64_t index, uint32_t* out, int capacity);

/* One whole-state row of the prepared state, each cell or pair pass launched
 * on its own and timed with HIP events (no graph, no height caps): ms_out[i] and
 * kind_out[i] for cell p = W-1-i, i < W-1 (kind 0 = cell_kernel, 1 =
 * cell_sorted_kernel, 2 = a pair pass over cells p and p-1, whose time it
 * carries, 3 = the second cell of a pair pass, time 0; a cell with several
 * launches, k > 16, sums them).
 * `capacity` slots each; fewer than W-1 is refused.  Advances the state by
 * one row.  Used by the benchmark for per-kernel rooflines; no reference
 * counterpart. */
ofp_status ofp_time_cells(ofp_ctx* ctx, float* ms_out, int* kind_out, int capacity);

/* Pair passes (default) or one launch per cell for the whole-state rows
 * of the context's own solves (ofp_solve, ofp_prepare + ofp_row_sweep,
 * ofp_apply_cells, ofp_time_cells).  A pair pass applies cells p and p-1 in
 * place, in up to five launches (one per window height change that has
 * classes; the three that read other classes side by side, then the other
 * two), so every word is read and written once per two cells instead of once
 * per cell.  Height-capped cells still run one launch per cell.  32-bit and
 * 64-bit ranks (n >= 23) alike; the static solve (ofp_solve_static) runs the
 * launches one after another.  Results do not depend on it.
 * mode: 0 = off, 1 = on (the default), 2..16 (even) = on with at most that
 * many moduli per launch (default 16).  Walking down a row, cells p and p-1 run
 * as one pair pass when bit p of cell_mask is set (~0u: everywhere), else cell
 * p runs alone; 0 = the library's default (pairs except over cells 4 and 3;
 * the autotuner sets it per size).  Bits 8..9
 * of mode select a kernel variant for 14 or 16 moduli per launch (tuning: bit 8
 * = closure sums gathered before the class's words are loaded, bit 9 = a
 * register budget for three workgroups per CU); results do not depend on it.
 * Other mode bits are refused (group sweeps: ofp_set_group_launch).
 * Reference analogue: the reference's fused two-cell group kernels
 * (row6crt20.cu:1033-1154), which run in group scratch. */
ofp_status ofp_set_pair_passes(ofp_ctx* ctx, int mode, uint32_t cell_mask);

/* How a group sweep (ofp_group_sweep) launches its steps.  streams: 1..4
 * (0 = the default, 3) streams that each step's group launches are dealt to,
 * by work; the groups of one step are independent, so their launches run side
 * by side, and every step still starts after the previous step of every group
 * has finished.  pair_mask: walking down the half, cells p and p-1 run as one
 * pair pass per group (the group's own class tables) where bit p is set, else
 * cell p runs alone; 0 = no pair passes, -1 = the library's default (pairs
 * from cell 7 up, cells 6..1 alone).  Results depend on neither.  No reference
 * counterpart for the streams (the reference runs one group kernel at a time
 * per GPU); the pair passes are the analogue of its fused two-cell group
 * kernels (row6crt20.cu:1033-1154). */
ofp_status ofp_set_group_launch(ofp_ctx* ctx, int streams, long long pair_mask);

/* Index width of the cell kernels: 0 = by n (32-bit ranks for n <= 22,
 * 64-bit above), 64 = 64-bit ranks at every n.  Results do not depend on it;
 * lets tests run the 64-bit kernels at sizes the CPU oracle checks.  No
 * reference counterpart. */
ofp_status ofp_set_index_width(ofp_ctx* ctx, int bits);

/* Kernel launch tuning (threads per workgroup, workgroups per CU); set by
 * the autotuner, 0 = default.  Reference analogue: GRIDFP_FRONTIER_THREADS /
 * GRIDFP_FRONTIER_BLOCKS (row6crt20.cu:1182,1192,1306-1307). */
ofp_status ofp_set_launch(ofp_ctx* ctx, int threads, int blocks_per_cu);

/* Cell-kernel variant for the state-major layout: 0 = an x=N thread loads its
 * own orbit values before scanning for closure predecessors, 1 = after (fewer
 * registers live during the scans), 2 = one branch-free orbit path for every
 * kind of blocked word (all lanes' loads issued before the scans; the
 * default), 3 = as 2 with the orbit results stored before the scans (fewer live
 * registers), 4 = as 2 with the sorted assignment on every cell (see
 * ofp_set_sorted_cells; shard sweeps run 2).  Set by the autotuner; no
 * reference counterpart. */
ofp_status ofp_set_variant(ofp_ctx* ctx, int variant);

/* Work distribution over the 8 XCDs: the blocked words are cut into chunks of
 * one workgroup and dealt to the XCDs in pieces of 2^log2_chunks chunks,
 * round-robin (default 5; a launch with fewer than 64 such pieces uses smaller
 * ones); -1 gives each XCD one contiguous eighth.  Set by the autotuner; no
 * reference counterpart. */
ofp_status ofp_set_xcd_pieces(ofp_ctx* ctx, int log2_chunks);

/* Full-state cells p >= from_cell (state-major layout) run the sorted kernel
 * (variant 4): each workgroup decodes its 256 blocked words, then deals them
 * back with the x=N words first, so that a wave holds one kind of word and its
 * orbit loads and stores are contiguous.  0 = never; -1 (default) = from cell 6
 * for launches carrying 8 or more moduli (HBM-bound), never otherwise (bound by
 * index math, where the barriers cost more than the contiguity gains).
 * Results do not depend on it.  Set by the autotuner; no reference counterpart. */
ofp_status ofp_set_sorted_cells(ofp_ctx* ctx, int from_cell);

/* Height caps.  In a sweep from the initial state (ofp_solve, ofp_prepare +
 * ofp_row_sweep, jobs) the cells of the first rows only touch words of small
 * height (every other word is still zero), and in the last two rows only such
 * words can still reach the answer word; a capped cell enumerates just those
 * blocked words.  A cell is capped when it then sweeps at most max_percent % of
 * the blocked words (default, and -1: 75 at one modulus, 85 with more; 0 turns
 * caps off).  The answer is the same;
 * other main words read after the last two rows are not final.  Reference
 * analogue: the row-6 prefix initialiser and the mirrored last row
 * (row6crt20.cu:1434-1688,1787-1814,935-952), which likewise skip the first
 * and last rows' full sweeps. */
ofp_status ofp_set_height_caps(ofp_ctx* ctx, int max_percent);

/* The height cap cell p of row `row` of an n-solve runs with at max_percent
 * (0: the whole state), and the number of blocked words it sweeps.
 * Host-only, no GPU.  Returns -1 for n out of range. */
int ofp_height_cap(int n, int row, int p, int max_percent, uint64_t* capped_blocked);

/* Moduli per launch on the low cells: cells p <= low_cells advance at most
 * low_kc moduli per launch (more launches, fewer registers per thread); the
 * other cells advance up to 16.  Default low_cells = 0.  Set by the autotuner;
 * no reference counterpart. */
ofp_status ofp_set_low_cells(ofp_ctx* ctx, int low_cells, int low_kc);

/* State layout for the next prepare/solve: 0 = tiles of 64 states (one
 * modulus of 64 consecutive states is 256 contiguous bytes), 1 = state-major
 * (the k residues of a state contiguous, stride padded to even).  No
 * reference counterpart (the reference runs one modulus at a time); set by
 * the autotuner.  Invalidates the prepared state. */
ofp_status ofp_set_layout(ofp_ctx* ctx, int layout);

/* ---- jobs: device-resident solves that run concurrently (serving) ----
 * A job owns its own state (main + blocked, k moduli interleaved) and its own
 * stream; row sweeps are enqueued asynchronously and the host polls for
 * completion, so several jobs' kernels overlap on the GPU while the host
 * admits and retires others.  No reference counterpart (the reference solves
 * one request per process); used by the serving loop (ofp/serving.py). */
typedef struct ofp_job ofp_job;
typedef struct ofp_job_info {
    int n, k, rows_total, rows_enqueued;
    int idle;                      /* 1 when all enqueued work has finished */
    uint64_t state_bytes;          /* device bytes held by the job */
    uint64_t transitions_per_row;
} ofp_job_info;
/* Allocate and initialise the state; fails with OFP_E_MEMORY if it does not fit. */
ofp_status ofp_job_create(ofp_ctx* ctx, int n, const uint32_t* moduli, int k, ofp_job** out);
/* Enqueue `rows` more row sweeps (at most W in total); returns immediately. */
ofp_status ofp_job_enqueue(ofp_job* job, int rows);
ofp_status ofp_job_info_get(ofp_job* job, ofp_job_info* info);
/* residues_out[j] = A007764(n) mod moduli[j]; waits for the job's stream. */
ofp_status ofp_job_result(ofp_job* job, uint32_t* residues_out);
void ofp_job_destroy(ofp_job* job);

/* ---- shard sweep: one GPU's share of the low cells of a row ----
 * The share is slice `slice` of `slices` equal ranges of the blocked index
 * space, widened to whole prefix classes (a prefix class = all words with the
 * same content at the positions above the low cells).  For cells
 * p = low_cells..1 every word such a share writes is in its own class, so the
 * share's main words form one contiguous window and its blocked words one
 * contiguous range; main words of other shards (read only by closures that
 * flip a position above the low cells) read as a seeded synthetic residue.
 * The window is initialised with the same synthetic residues ("input residues
 * seeded uniform in [0,p)"), the cells are swept in place on the device, and
 * the kernel time is returned.  Reference analogue: one GPU's owner-local
 * LOW window in the 8-GPU bucket layout (docs/b300-bucket-snake.md:7-20). */
typedef struct ofp_shard_spec {
    int n;
    int k;                     /* 1, 2, 4 or 8 moduli, interleaved */
    const uint32_t* moduli;
    int slices, slice;         /* this shard = slice of slices */
    int low_cells;             /* cells low_cells..1 (0 = (n+1)/2) */
    uint64_t seed;
} ofp_shard_spec;
typedef struct ofp_shard_info {
    uint64_t first_blocked, blocked_count;   /* resident blocked words */
    uint64_t first_main, main_count;         /* resident main window */
    uint64_t state_bytes;                    /* device bytes of the window */
    int low_cells, launches;
    double kernel_s;                         /* all cells, HIP events */
    uint64_t stray_stores;                   /* stores outside the window (0 by construction) */
} ofp_shard_info;
/* Window of a shard (no device work; ctx may be NULL). */
ofp_status ofp_shard_plan(ofp_ctx* ctx, const ofp_shard_spec* spec, ofp_shard_info* info);
/* main_out (main_count*k) / blocked_out (blocked_count*k): optional host copies
 * of the window after the sweep, state-major.  Shard and group sweeps keep
 * their window buffer in the context from call to call (streaming windows of
 * one size re-touches no pages and frees nothing); ofp_destroy frees it, and
 * so does a full-state sweep (ofp_solve, ofp_prepare, ...) before sizing its
 * own state. */
ofp_status ofp_shard_sweep(ofp_ctx* ctx, const ofp_shard_spec* spec, ofp_shard_info* info,
                           uint32_t* main_out, uint32_t* blocked_out);

/* ---- group sweep: one half of a row over occupancy groups ----
 * A row of width W = n+1 is split at `low` into its high cells W-1..low+1 and
 * its low cells low..1.  Each half rewrites only its own positions and turns a
 * closure partner's R into L or back, which never changes whether a position
 * is occupied (N or not), so the words with one occupancy pattern of the
 * positions outside the half (a group) are closed under it:
 *   high = 0: LOW groups, occupancy of positions 0..low-1, cells W-1..low+1;
 *             the first cell reads no blocked input (a row's start), the last
 *             leaves the window's blocked words (blocked_io is output);
 *   high = 1: HIGH groups, occupancy of positions low+1..W-1, cells low..1;
 *             the first cell reads the window's blocked words (blocked_io is
 *             input: the high half's output), after cell 1 none are left.
 * A window is the groups first_group..first_group+groups-1 (masks in
 * increasing order; bit i of a mask is the first fixed position + i).  Each
 * group is stored contiguously with its words in global rank order (32-bit
 * ranks local to the group: every group of the window must hold fewer than
 * 2^32 words); the window's main array is the groups' main words one group
 * after the other, its blocked array likewise.  The cells run in place, two
 * per pass where group pair passes are on (ofp_set_group_launch: the pair
 * pass over each group with the group's own class tables; cell_ms then holds a
 * pass's time at its first cell and 0 at its second), else one per launch.
 * Reference analogue: one GPU's owner-local LOW and HIGH windows of the 8-GPU
 * bucket layout (docs/b300-bucket-snake.md:7-20; docs/b300-hbm32.md:251-270:
 * W = 28, LOW14 / HIGH13, 16,384 and 8,192 groups). */
typedef struct ofp_group_spec {
    int n;
    int k;                     /* 1..16 moduli, interleaved */
    const uint32_t* moduli;
    int low;                   /* split: 1..W-2 (0 = W/2); at most 24 fixed positions */
    int high;                  /* 0: LOW groups / high cells, 1: HIGH groups / low cells */
    uint32_t first_group, groups;
    uint64_t seed;             /* seeded inputs (uniform in [0, p)) where no input is given */
} ofp_group_spec;
typedef struct ofp_group_info {
    uint64_t main_count, blocked_count;  /* words of the window */
    uint64_t largest_group;              /* main + blocked words of its largest group */
    uint64_t state_bytes;                /* device bytes of the window */
    int low, cells, launches;            /* cells in the half: W-1-low or low */
    double kernel_s;                     /* all cells, HIP events */
    float cell_ms[OFP_MAX_N];            /* cell_ms[i]: the i-th cell of the half, top down */
    int steps;                           /* passes over the window: pair passes + single cells */
} ofp_group_info;
/* Window totals into info, and per-group counts into main_counts /
 * blocked_counts (groups entries each; either may be NULL).  No device work;
 * ctx may be NULL. */
ofp_status ofp_group_plan(ofp_ctx* ctx, const ofp_group_spec* spec, ofp_group_info* info,
                          uint64_t* main_counts, uint64_t* blocked_counts);
/* main_io: NULL for seeded inputs, else main_count*k residues (state-major,
 * window order) that are replaced by the result.  blocked_io: NULL or
 * blocked_count*k residues, written (high = 0) or read (high = 1; NULL: seeded). */
ofp_status ofp_group_sweep(ofp_ctx* ctx, const ofp_group_spec* spec, ofp_group_info* info,
                           uint32_t* main_io, uint32_t* blocked_io);

/* ---- exchange: a LOW window's words to the shares of their HIGH groups ----
 * Between the two halves of a row every word moves from its LOW group to its
 * HIGH group (main words: occupancy of main positions low+1..W-1; blocked
 * words: of blocked positions low..W-2).  With the HIGH groups owned by
 * `shares` GPUs (runs of consecutive masks, cut at s/shares of the main +
 * blocked words, s = 1..shares-1), this writes every word of the LOW window
 * that the last ofp_group_sweep of this context left on the device (spec must
 * name that window, high = 0) into a staging bucket per destination share, as
 * a record (index of the word in that share's HIGH placement -- the window
 * layout ofp_group_sweep takes for the share's masks -- and its residues).
 * Records keep the window order inside a bucket; bucket s follows buckets
 * 0..s-1.  On one GPU the buckets of the other shares stand for the send
 * buffers of the per-row all-to-all.  Needs fewer than 17 fixed HIGH
 * positions.  Reference analogue: the per-row 8x8 bucket transpose between the
 * H-major and L-major placements (docs/b300-bucket-snake.md:7-20), whose
 * owner pairs B[a, b] are this build's (HIGH share, LOW share) pairs. */
#define OFP_MAX_SHARES 16
typedef struct ofp_exchange_info {
    uint64_t main_count, blocked_count;           /* words of the window */
    int shares;
    uint32_t share_first[OFP_MAX_SHARES + 1];     /* share s: HIGH masks share_first[s]..share_first[s+1]-1 */
    uint64_t main_records[OFP_MAX_SHARES];        /* records per destination share: main words */
    uint64_t blocked_records[OFP_MAX_SHARES];     /*                                 blocked words */
    uint64_t staging_bytes;                       /* device bytes of the staging buckets */
    uint64_t moved_bytes;                         /* residues read + records written */
    double kernel_s;                              /* count, scan and write launches, HIP events */
} ofp_exchange_info;
/* main_index / blocked_index (main_count / blocked_count entries) and
 * main_res / blocked_res (x k residues): optional host copies of the buckets,
 * all buckets in order. */
ofp_status ofp_group_exchange(ofp_ctx* ctx, const ofp_group_spec* spec, int shares, ofp_exchange_info* info,
                              uint64_t* main_index, uint32_t* main_res, uint64_t* blocked_index,
                              uint32_t* blocked_res);

/* Counts only, no device needed: main / blocked word counts for n. */
ofp_status ofp_state_counts(int n, uint64_t* main_states, uint64_t* blocked_states);

const char* ofp_last_error(const ofp_ctx* ctx);
const char* ofp_version(void);
void ofp_destroy(ofp_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* OFP_H */
