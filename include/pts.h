/* C ABI of the MI355X-native heads-up DCFR postflop solver (libpts.so).
 *
 * Two layers:
 *  - pts_engine_*: the drop-in for the reference's GPU backend. It takes the
 *    caller's game tree, hand lists and regret / strategy-sum storage and runs
 *    DCFR iterations, best responses and subtree evaluations in HIP kernels on
 *    an MI355X (gfx950). It replaces crates/solver/src/gpu/{mod,plan,engine}.rs
 *    (`Gpu`, `Plan::build_with`, `Engine`) and the CUDA kernels in
 *    crates/solver/cuda/.
 *  - pts_solver_*: the solver crate's public API (crates/solver/src/lib.rs:
 *    `Solver::{estimate,new,solve,query,ranges_after}`) over that engine, for
 *    callers without a tree builder of their own.
 * All functions return a pts_status; pts_last_error() describes the last
 * failure on the calling thread. No function aborts the process.
 *
 * Threads: a pts_solver may be used from several threads; its calls that
 * run the GPU (solve, iterate, exploitability, root values, storage,
 * queries) take one lock per solver, so queries from reader threads wait
 * for a solve in progress (the reference queries after solves from other
 * threads, crates/trainer/src/study.rs:995-1051). A pts_engine, like the
 * reference's gpu::Engine, is used from one thread at a time.
 */
#ifndef PTS_H
#define PTS_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* crates/solver/src/gpu/mod.rs:86-136 (GpuError), as status codes */
typedef enum {
  PTS_OK = 0,
  PTS_DRIVER_MISSING = 1,
  PTS_DRIVER_TOO_OLD = 2,
  PTS_NO_DEVICE = 3,
  PTS_DRIVER = 4,
  PTS_INVALID_INPUT = 5,
  PTS_OUT_OF_MEMORY = 6,
  PTS_CANCELLED = 7,
} pts_status;

/* crates/solver/src/cfr.rs:248-256 */
typedef enum { PTS_MODE_CFR = 0, PTS_MODE_BEST_RESPONSE = 1, PTS_MODE_AVERAGE = 2 } pts_mode;

/* crates/solver/src/lib.rs:57-71 */
typedef enum { PTS_BACKEND_CPU = 0, PTS_BACKEND_GPU = 1, PTS_BACKEND_AUTO = 2 } pts_backend;

/* How terminal sums are accumulated (crates/solver/src/gpu/mod.rs:51-63):
 * FAST adds the long strength sweeps as f64 block scans (the values differ
 * from the CPU solve's by rare one-rounding-step f32 differences); EXACT keeps
 * the CPU's sequential f64 order and is bit-identical to it. */
typedef enum { PTS_PRECISION_FAST = 0, PTS_PRECISION_EXACT = 1 } pts_precision;

typedef struct pts_engine pts_engine;
typedef struct pts_solver pts_solver;

/* sizeof of the ABI structs, for binders to check their layouts:
 * 0 pts_config_c, 1 pts_tree_desc, 2 pts_progress_c, 3 pts_step_c, 4 pts_stats_c. */
size_t pts_abi_sizeof(int what);
/* Text of the last error on this thread; returns its full length. */
size_t pts_last_error(char* buf, size_t len);
/* The last failure's numbers (thread-local; zeros when it has none): the
 * hipError_t of a PTS_DRIVER (or PTS_NO_DEVICE) status, and the bytes
 * needed and available of a PTS_OUT_OF_MEMORY status, the fields of the
 * reference's GpuError::Driver and GpuError::OutOfMemory
 * (crates/solver/src/gpu/mod.rs:86-136). */
void pts_last_error_detail(int32_t* hip_error, uint64_t* needed, uint64_t* available);
/* evaluate_cards (crates/holdem/src/eval.rs:142-198) over `count` hands of
 * `ncards` (5 to 7) distinct cards each (card = rank * 4 + suit), as the tree
 * builder's strength tables use it: out[i] = category << 26 | payload. */
pts_status pts_evaluate(const uint8_t* cards, uint32_t ncards, uint64_t count, uint32_t* out);
/* Number of usable GPUs (crates/solver/src/gpu/mod.rs:242-282 opens device 0). */
pts_status pts_device_count(int* out);
/* Whether `backend` puts a solve rooted on `street` (1 flop, 2 turn, 3 river)
 * on the GPU (crates/solver/src/lib.rs:497-503). Auto differs from the
 * reference's: on MI355X river roots solve faster on the GPU too (build
 * included), so Auto means the GPU for every street. */
int pts_backend_wants_gpu(int backend, int street);

/* ------------------------------------------------------------------ engine */

/* The caller's game tree (the arena of crates/solver/src/tree.rs:154-245),
 * hand lists (crates/solver/src/hands.rs:14-30) and strength-table boards. */
typedef struct {
  uint64_t num_nodes;
  const uint8_t* kind;          /* 0/1 decision of player 0/1, 2 chance, 3/4 fold by player 0/1, 5 showdown */
  const uint8_t* street;        /* 1 flop, 2 turn, 3 river */
  const uint16_t* num_children;
  const uint32_t* first_child;  /* children are contiguous */
  const int32_t* invest;        /* 2 per node: chips each player put in since the root */
  const uint64_t* offset;       /* decision: storage offset; showdown: board index */
  const float* rake;            /* terminals */
  const float* chance_factor;   /* chance nodes */
  const int8_t* card;           /* chance children: dealt card; else -1 */
  uint32_t num_boards;
  const int32_t* boards;        /* 5 cards per showdown board */
  uint32_t num_hands[2];
  const uint8_t* hand_cards[2]; /* 2 per hand, higher card index first */
  const float* hand_weights[2];
  const uint32_t* hand_same[2]; /* index of the identical combo in the other list, or 0xFFFFFFFF */
  double half_pot;              /* root pot / 2 */
  double num_combinations;      /* total weight of compatible hand pairs */
  uint64_t storage_len;         /* regret / strategy-sum cells */
  int32_t batch_river;          /* river subtrees use the batched evaluators' rounding (the reference default) */
  int32_t _pad;
} pts_tree_desc;

/* Gpu::new + Plan::build_with + Engine::new (gpu/mod.rs:244, plan.rs:270,
 * engine.rs:292): validates and flattens the tree, uploads it and the storage
 * (regrets / sums may be null: zeros). scratch_budget (Engine::scratch_budget
 * and Scratch, engine.rs:286, plan.rs:270) caps the bytes of reach and value
 * rows; 0 uses the HBM left after the storage. The subtrees below the root
 * street's chance nodes run in groups that fit it. */
pts_status pts_engine_create(int device, const pts_tree_desc* tree, const float* regrets, const float* sums, int precision,
                             uint64_t scratch_budget, pts_engine** out);
/* The same engine over 16-bit storage (storage.rs:87-106; the reference's
 * HostStorage::Compressed, gpu/engine.rs:292-343): storage_len i16 regrets
 * and u16 strategy sums; slot[i] is decision node i's scale slot (tree.rs:
 * 164-165, Node::slot; ignored for other nodes), num_slots regret scales and
 * num_slots strategy-sum scales; the four storage arrays may all be null
 * (zeros). */
pts_status pts_engine_create_q(int device, const pts_tree_desc* tree, const uint32_t* slot, uint32_t num_slots,
                               const int16_t* regrets, const uint16_t* sums, const float* regret_scale,
                               const float* strat_scale, int precision, uint64_t scratch_budget, pts_engine** out);
/* engine.rs:347-382 for 16-bit storage: the integers and the scales. */
pts_status pts_engine_download_q(pts_engine* e, int16_t* regrets, uint16_t* sums, float* regret_scale, float* strat_scale);
/* Engine::cfr_iteration (engine.rs:386-392): player 0's update, then player 1's,
 * with the discounts of cfr.rs:83-107 computed by the caller. */
pts_status pts_engine_iterate(pts_engine* e, float pos, float neg, float strat, int reach_weighted);
/* One traversal of `player` under `mode` (CFR mode updates storage); the
 * player's root values (FP64 for best response / average). */
pts_status pts_engine_traverse(pts_engine* e, int player, int mode, float pos, float neg, float strat, int reach_weighted,
                               double* out, size_t n);
/* Engine::root_values (engine.rs:406): expected payoffs in chips, traversals in FP64. */
pts_status pts_engine_root_values(pts_engine* e, int mode, double out[2]);
/* Engine::exploitability (engine.rs:417): FP64 best responses of both players. */
pts_status pts_engine_exploitability(pts_engine* e, int raked, double* out);
/* Engine::download (engine.rs:347): regrets and strategy sums to the host. */
pts_status pts_engine_download(pts_engine* e, float* regrets, float* sums, uint64_t len);
pts_status pts_engine_upload(pts_engine* e, const float* regrets, const float* sums, uint64_t len);
/* Average-strategy values of the children of decision node `node` (player
 * `player`) given the opponent's reach there (the traversals of
 * Solver::query, lib.rs:854-863): out[action * n_player + hand]. */
pts_status pts_engine_subtree_values(pts_engine* e, uint32_t node, int player, const float* opp_reach, size_t nq, float* out,
                                     size_t len);
/* One DCFR iteration launched kernel by kernel with HIP events around every
 * launch on the engine's stream: summed milliseconds and launch counts per
 * kernel class [0 begin, 1 down (reach push-down + strategy sums), 2 terminal
 * (fold / showdown values), 3 up (value pull-up + regret update)], and the
 * wall time from the first event to the last. */
pts_status pts_engine_profile_iteration(pts_engine* e, float pos, float neg, float strat, int reach_weighted, double ms[4],
                                        uint32_t launches[4], double* wall_ms);
/* Bytes of HBM the engine holds. */
uint64_t pts_engine_device_bytes(pts_engine* e);
/* Chunk groups one traversal runs in (f64 = 0: DCFR rows, 1: FP64 best-response rows). */
uint32_t pts_engine_num_groups(pts_engine* e, int f64);
/* Segments whose DCFR up walk runs in a kernel specialised to the segment's
 * shape, compiled at creation: in a plan with at least 8192 segments that
 * have a walk, for every shape of at least 64 of them (none in a smaller
 * plan). PTS_JIT_MIN_SEGMENTS=n sets the per-shape count and drops the
 * plan-size rule; PTS_JIT=0 turns them off. No reference counterpart: the
 * reference's walk is per level (gpu/engine.rs:600-749). */
uint64_t pts_engine_jit_segments(pts_engine* e);
void pts_engine_destroy(pts_engine* e);

/* ------------------------------------------------------------------ solver */

typedef struct {
  int32_t kind[8]; /* 0 pot fraction, 1 multiple of the current bet, 2 all-in (crates/holdem/src/menu.rs:14-24) */
  double value[8];
  int32_t count;
  int32_t _pad;
} pts_sizes_c;

typedef struct { /* crates/holdem/src/menu.rs:100-119 */
  pts_sizes_c bet[3]; /* flop, turn, river */
  pts_sizes_c raise[3];
  uint32_t max_raises;
  uint32_t _pad;
  double add_allin_threshold, force_allin_threshold, merge_threshold;
} pts_menu_c;

typedef struct { /* crates/solver/src/lib.rs:73-105 */
  int32_t board[5];
  int32_t board_len;
  float ranges[2][1326];
  int32_t street; /* root betting state (crates/solver/src/tree.rs:19-34) */
  int32_t to_act;
  int64_t pot, commit[2], stack[2], current_bet, last_raise_inc;
  uint32_t num_raises;
  int32_t needs[2];
  int32_t has_future_menu;
  pts_menu_c menu, future_menu;
  int64_t big_blind;
  double rake_rate, rake_cap;
  uint64_t max_memory_bytes;
  float prune;        /* must be 0 */
  int32_t storage_mode; /* storage.rs:74-81: 0 f32, 1 16-bit ("compressed"), 2 auto (16-bit above 1 GB of f32);
                           batched and depth-limited solves take f32 only */
  int32_t isomorphism;  /* suit isomorphism (iso.rs) */
  int32_t beta_some;
  double alpha, beta, gamma; /* crates/solver/src/cfr.rs:47-63 */
  int32_t restart_average, unweighted_average;
  int32_t backend;
  /* tree.rs:128-136 ChanceMode: 0 Exact; 1 Sampled { turn, river, seed }: a
   * stratified sample of sample_turn turn cards and sample_river river cards
   * per branch (sampling.rs; all-in runouts stay exact), drawn from rand
   * 0.9's StdRng seeded with sample_seed */
  int32_t chance_mode;
  uint32_t sample_turn, sample_river;
  uint64_t sample_seed;
} pts_config_c;

typedef struct { uint64_t player_nodes, chance_nodes, terminal_nodes, storage_cells, memory_bytes; } pts_stats_c;

typedef struct { /* crates/solver/src/lib.rs:174-184 */
  uint32_t iteration, max_iterations;
  double exploitability, exploitability_pct;
  int32_t finished, _pad;
} pts_progress_c;

typedef struct { /* crates/solver/src/lib.rs:212-217: kind 0 action (act_kind 0 fold, 1 check, 2 call, 3 raise to `to`), 1 card */
  int32_t kind, act_kind;
  int64_t to;
  int32_t card, _pad;
} pts_step_c;

/* Solver::estimate (lib.rs:355-362) */
pts_status pts_solver_estimate(const pts_config_c* cfg, pts_stats_c* out);
/* Solver::new (lib.rs:365-396) plus the engine (Solver::use_gpu,
 * lib.rs:656-660): the tree is built on the host and lives in HBM with its
 * storage. device -1 builds and validates on the host only. */
pts_status pts_solver_new(const pts_config_c* cfg, int device, int precision, uint64_t scratch_budget, pts_solver** out);
/* Chunk groups of the solver's engine (as pts_engine_num_groups). */
uint32_t pts_solver_num_groups(pts_solver* s, int f64);
/* Shape-specialised walks of the solver's engine (as pts_engine_jit_segments). */
uint64_t pts_solver_jit_segments(pts_solver* s);
/* Host-only solvers (device -1): the chunk groups a scratch budget gives and
 * the scratch bytes they need, after checking that every slot and row a
 * grouped traversal touches maps inside that scratch (Plan::build_with's
 * Scratch sizing, plan.rs:270). */
pts_status pts_solver_plan_groups(pts_solver* s, uint64_t scratch_budget, int f64, uint32_t* num_groups,
                                  uint64_t* scratch_bytes);
void pts_solver_free(pts_solver* s);
void pts_solver_stats(pts_solver* s, pts_stats_c* out);
/* Solver::solve (lib.rs:475-642). `history` (may be null) receives the
 * progress of every exploitability check. */
pts_status pts_solver_solve(pts_solver* s, uint32_t max_iterations, double target_pct, uint32_t check_every,
                            pts_progress_c* out, pts_progress_c* history, int32_t history_cap, int32_t* history_len);
/* Called after every exploitability check with the progress so far
 * (crates/solver/src/lib.rs:614-630, `on_progress`); returning 0 stops the
 * solve. */
typedef int32_t (*pts_progress_fn)(const pts_progress_c* progress, void* user);
/* Solver::solve_cancellable (lib.rs:480-494): pts_solver_solve that reads
 * *cancel (null: never set) with a relaxed atomic load before every
 * iteration and stops within one iteration once it is nonzero (the layout of
 * Rust's AtomicBool: one byte), and calls on_progress (may be null) after
 * every check. A solve stopped either way returns PTS_CANCELLED with `out`
 * holding the progress of its last check and finished = 1; the storage keeps
 * every completed iteration, so a later solve continues from there. */
pts_status pts_solver_solve_cancellable(pts_solver* s, uint32_t max_iterations, double target_pct, uint32_t check_every,
                                        const uint8_t* cancel, pts_progress_fn on_progress, void* user,
                                        pts_progress_c* out);
/* The solve loop the two calls above share, for engine-layer callers that
 * keep their own solver state (Solver::run_solve, lib.rs:576-642): runs
 * iterate(user, t) for t = *iterations + 1, ... (a pts_status, 0 for OK;
 * *iterations counts it first), measures the exploitability in chips with
 * measure(user, &x) on the reference's adaptive schedule (next_check_gap,
 * lib.rs:155-172: every check_every iterations until the trend predicts the
 * target sooner) and at the last iteration, and stops at the first check at
 * or below target_pct of `pot`, at max_iterations, or when cancelled as
 * pts_solver_solve_cancellable is. *exploitability holds the last check. A
 * status other than PTS_OK from iterate or measure ends the loop and is
 * returned. */
typedef int32_t (*pts_iterate_fn)(void* user, uint32_t iteration);
typedef int32_t (*pts_measure_fn)(void* user, double* exploitability);
pts_status pts_run_solve(uint32_t* iterations, double* exploitability, double pot, uint32_t max_iterations,
                         double target_pct, uint32_t check_every, const uint8_t* cancel, pts_iterate_fn iterate,
                         pts_measure_fn measure, pts_progress_fn on_progress, void* user, pts_progress_c* out);
/* `count` DCFR iterations without exploitability checks. */
pts_status pts_solver_iterate(pts_solver* s, uint32_t count);
pts_status pts_solver_exploitability(pts_solver* s, double* out);
pts_status pts_solver_root_values(pts_solver* s, int mode, double out[2]);
uint32_t pts_solver_iterations(pts_solver* s);
void pts_solver_set_iterations(pts_solver* s, uint32_t t);
uint64_t pts_solver_storage_len(pts_solver* s);
int32_t pts_solver_num_hands(pts_solver* s, int player);
/* f32 storage as stored; 16-bit storage decoded, q * (scale / MAX) (storage.rs:209-246).
 * Setting f32 values is refused for 16-bit storage. */
pts_status pts_solver_get_storage(pts_solver* s, float* regrets, float* sums);
pts_status pts_solver_set_storage(pts_solver* s, const float* regrets, const float* sums);
/* 16-bit storage (storage.rs:87-106; the reference's HostStorage::Compressed,
 * gpu/mod.rs, which its engine uploads and downloads): storage_len i16
 * regrets and u16 strategy sums, and one regret scale and one strategy-sum
 * scale per player node (player_nodes of pts_solver_stats, in the tree
 * builder's order, tree.rs:463-469). */
int32_t pts_solver_is_compressed(pts_solver* s);
pts_status pts_solver_get_storage_q(pts_solver* s, int16_t* regrets, uint16_t* sums, float* regret_scale, float* strat_scale);
pts_status pts_solver_set_storage_q(pts_solver* s, const int16_t* regrets, const uint16_t* sums, const float* regret_scale,
                                    const float* strat_scale);
/* The next DCFR iteration (advancing the iteration count), profiled as
 * pts_engine_profile_iteration. */
pts_status pts_solver_profile_iteration(pts_solver* s, double ms[4], uint32_t launches[4], double* wall_ms);
/* The arena (same layout as pts_tree_desc; arrays sized num_nodes). */
uint64_t pts_solver_num_nodes(pts_solver* s);
void pts_solver_nodes(pts_solver* s, uint8_t* kind, uint8_t* street, uint16_t* num_children, uint32_t* first_child,
                      int32_t* invest, uint64_t* offset, float* rake, float* chance_factor, int8_t* card, int32_t* edge_act_kind,
                      int64_t* edge_to);
/* Solver::query (lib.rs:842-899): strategy and per-action EVs at the
 * decision node reached by `path`. */
pts_status pts_solver_query(pts_solver* s, const pts_step_c* path, int32_t npath, int32_t* player, int32_t* na, int32_t* n,
                            int32_t* act_kinds, int64_t* act_tos, uint16_t* combos, float* strategy, float* ev, float* reach,
                            double reach_fraction[2], int64_t* pot, int64_t invest[2], int32_t cap_actions, int32_t cap_hands);
/* Solver::ranges_after (lib.rs:903-912) */
pts_status pts_solver_ranges_after(pts_solver* s, const pts_step_c* path, int32_t npath, float* out0, float* out1);

/* ------------------------------------------------------------------ batch */
/* Many independent spots solved together (no counterpart in the reference,
 * which solves one spot per Solver): every spot's tree is built as
 * pts_solver_new builds it, the trees are merged into one plan (segments of
 * all spots by depth), and each DCFR iteration of every active spot is one
 * hipGraph launch. The spots share one set of DCFR parameters and one
 * iteration count; a spot's storage evolves exactly as it would alone. */
typedef struct pts_batch pts_batch;
pts_status pts_batch_new(const pts_config_c* cfgs, uint32_t num_spots, int device, int precision, pts_batch** out);
void pts_batch_free(pts_batch* b);
uint32_t pts_batch_num_spots(pts_batch* b);
uint32_t pts_batch_iterations(pts_batch* b);
pts_status pts_batch_stats(pts_batch* b, uint32_t spot, pts_stats_c* out);
uint64_t pts_batch_storage_len(pts_batch* b, uint32_t spot);
pts_status pts_batch_get_storage(pts_batch* b, uint32_t spot, float* regrets, float* sums);
/* `count` DCFR iterations of every active spot. */
pts_status pts_batch_iterate(pts_batch* b, uint32_t count);
/* Exploitability of every spot in chips (FP64 best responses, raked as Solver). */
pts_status pts_batch_exploitability(pts_batch* b, double* out);
/* Iterates until every spot's exploitability is at most target_pct of its
 * pot or max_iterations; out[spot] is each spot's progress. Each spot keeps
 * the single solve's check schedule (lib.rs:155-172,594-642: every
 * check_every iterations until its own trend predicts the target sooner);
 * a check measures only the spots due, and a spot at its target is frozen
 * (its traversals skipped). */
pts_status pts_batch_solve(pts_batch* b, uint32_t max_iterations, double target_pct, uint32_t check_every,
                           pts_progress_c* out);
/* Autotuner: times `iterations` iterations for every tile choice (terminals
 * per terminal workgroup 1, 2, 4, 8, 16, 32, 64 or 256; opponent hands per
 * down workgroup 128/256/512/1024), keeps the fastest (best[0], best[1],
 * best_ms per iteration) and restarts the solve from zero storage. */
pts_status pts_batch_autotune(pts_batch* b, uint32_t iterations, uint32_t best[2], double* best_ms);
/* Sets the tiles: run_length 1..256 terminals per terminal workgroup (values
 * above 256 are clamped), down_threads 128/256/512/1024. The tiles change
 * launch shapes only, not the arithmetic. */
pts_status pts_batch_set_tiles(pts_batch* b, uint32_t run_length, uint32_t down_threads);
/* Every spot's per-hand root values (FP64 traversal under `mode`: best
 * response or average) in combo order: out[(spot * 2 + player) * 1326 +
 * combo], 0 for combos outside the range. Values are normalized as the
 * solver's (chips x reach / compatible weight). */
pts_status pts_batch_root_rows(pts_batch* b, int mode, double* out);
/* The spot's total weight of compatible hand pairs (the normalizer). */
double pts_batch_num_combinations(pts_batch* b, uint32_t spot);
pts_status pts_batch_profile_iteration(pts_batch* b, double ms[4], uint32_t launches[4], double* wall_ms);
/* Sizes of the merged plan; for a host-only batch (device -1, built and
 * checked without a GPU) also the scratch bytes of its DCFR and FP64 rows. */
pts_status pts_batch_host_info(pts_batch* b, uint64_t* num_nodes, uint64_t* storage_len, uint64_t scratch[2]);

/* ------------------------------------------------------------------ value network */
/* The counterfactual-value MLP of the depth-limited flop solve (no
 * counterpart in the reference): 2706 inputs (52-card board mask, pot and
 * stack as fractions of pot + stack, both ranges normalized to sum 1) ->
 * 4 x 1024 ReLU -> 2652 outputs (both players' expected values per combo, in
 * pot fractions). BF16 weights and activations, FP32 accumulation, hand-
 * written MFMA kernels; masked Huber loss, Adam. Rows per call: a positive
 * multiple of 64 up to the capacity. */
typedef struct pts_net pts_net;
uint64_t pts_net_num_params(void);
double pts_net_forward_flops_per_row(void);
/* He-normal weights from `seed`, zero biases. */
pts_status pts_net_create(int device, uint32_t batch_cap, uint64_t seed, pts_net** out);
void pts_net_destroy(pts_net* n);
/* Packed f32 parameters: W0 (1024 x 2706, row-major), b0, W1, b1, W2, b2, W3,
 * b3, W4 (2652 x 1024), b4. Setting them resets the Adam state. */
pts_status pts_net_set_params(pts_net* n, const float* params);
pts_status pts_net_get_params(pts_net* n, float* params);
/* Gradients of the last pts_net_gradients / pts_net_train_step, packed alike. */
pts_status pts_net_get_grads(pts_net* n, float* grads);
pts_status pts_net_forward(pts_net* n, const float* x, uint32_t rows, float* y);
/* pts_net_forward without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). x and y are device rows
 * (rows x 2706 and rows x 2652 f32), rows a positive multiple of 64 up to
 * 1024; the work is enqueued on `stream` (a hipStream_t; null: the
 * network's own stream) and not waited for. The working rows live in one
 * fixed device array of 20.8 MB per GPU (taken at pts_net_create), shared by
 * every call on that GPU: calls that overlap on one GPU race. */
pts_status pts_net_forward_noalloc(pts_net* n, const float* x, uint32_t rows, float* y, void* stream);
pts_status pts_net_gradients(pts_net* n, const float* x, const float* target, const float* mask, uint32_t rows, double* loss);
pts_status pts_net_train_step(pts_net* n, const float* x, const float* target, const float* mask, uint32_t rows, float lr,
                              double* loss);
pts_status pts_net_profile(pts_net* n, uint32_t rows, double* fwd_ms, double* train_ms);

/* ------------------------------------------------------------------ depth-limited solve */
/* A solver over the tree cut where the root street's betting ends with
 * chips behind (all-in lines keep their exact runouts); every cut point is
 * valued per dealt card by `net` (null: zero), on the GPU inside each
 * iteration. Returns a pts_solver; its exploitability is the depth-limited
 * game's. (New here: the reference has no value network.) */
pts_status pts_dl_new(const pts_config_c* cfg, int device, int precision, pts_net* net, pts_solver** out);  /* device -1: host only */
uint32_t pts_solver_num_leaves(pts_solver* s);
/* Copies src's root-street decision nodes (a depth-limited solver) into the
 * same action paths of dst (the full tree), scaling regrets and strategy
 * sums: the warm start of the exact solve that judges a value-network
 * solve. */
pts_status pts_solver_copy_root_street(pts_solver* dst, pts_solver* src, float regret_scale, float sum_scale,
                                       uint32_t* copied);
/* Pins dst's root-street decision nodes to src's average strategy (no
 * reference counterpart; the value-net judge's re-solve of the later streets
 * under the depth-limited root strategy): sums copied, regrets set to the
 * average strategy's probabilities so that regret matching plays it, and
 * those nodes frozen: DCFR iterations leave their regrets as written, so
 * both traversals of every iteration play the pinned strategy (the strategy
 * sums keep growing in the same proportions). The nodes stay frozen for the
 * solver's life. */
pts_status pts_solver_pin_root_street(pts_solver* dst, pts_solver* src, uint32_t* pinned);

#ifdef __cplusplus
}
#endif
#endif
