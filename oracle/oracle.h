/* C API of the CPU oracle (test infrastructure only: loaded by tests/, the
 * smoke entry point and the benchmark's cpu_baseline leg, never by the
 * product library). The config layout is the same as pts_config_c in
 * include/pts.h so one ctypes structure serves both. */
#ifndef PTS_ORACLE_H
#define PTS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t kind[8]; /* 0 pot fraction, 1 multiple of the current bet, 2 all-in */
  double value[8];
  int32_t count;
  int32_t _pad;
} orc_sizes_c;

typedef struct {
  orc_sizes_c bet[3];   /* flop, turn, river */
  orc_sizes_c raise[3];
  uint32_t max_raises;
  uint32_t _pad;
  double add_allin_threshold, force_allin_threshold, merge_threshold;
} orc_menu_c;

typedef struct {
  int32_t board[5];
  int32_t board_len;
  float ranges[2][1326];
  int32_t street; /* 1 flop, 2 turn, 3 river */
  int32_t to_act;
  int64_t pot, commit[2], stack[2], current_bet, last_raise_inc;
  uint32_t num_raises;
  int32_t needs[2];
  int32_t has_future_menu;
  orc_menu_c menu, future_menu;
  int64_t big_blind;
  double rake_rate, rake_cap;
  uint64_t max_memory_bytes;
  float prune;
  int32_t storage_mode; /* 0 f32, 1 16-bit, 2 16-bit above 1 GB of f32 */
  int32_t isomorphism;
  int32_t beta_some;
  double alpha, beta, gamma;
  int32_t restart_average, unweighted_average;
  int32_t backend; /* 0 cpu, 1 gpu, 2 auto */
  int32_t chance_mode; /* 0 exact, 1 sampled (tree.rs:128-136) */
  uint32_t sample_turn, sample_river;
  uint64_t sample_seed;
} orc_config_c;

typedef struct {
  uint64_t player_nodes, chance_nodes, terminal_nodes, storage_cells, memory_bytes;
} orc_stats_c;

typedef struct {
  uint32_t iteration, max_iterations;
  double exploitability, exploitability_pct;
  int32_t finished, _pad;
} orc_progress_c;

/* one step of a path: kind 0 action (act_kind 0 fold 1 check 2 call 3 raise, to), 1 card */
typedef struct {
  int32_t kind, act_kind;
  int64_t to;
  int32_t card, _pad;
} orc_step_c;

/* sizeof(orc_config_c) */
size_t orc_config_size(void);

/* primitives */
uint32_t orc_evaluate(const int32_t* cards, int32_t n);
/* evaluate_cards over `count` hands of `ncards` cards each */
void orc_evaluate_batch(const uint8_t* cards, int32_t ncards, uint64_t count, uint32_t* out);
/* The `count` ncards-card hands (cards ascending) from lexicographic rank
 * `first` on, over the 52-card deck; returns how many were written. */
uint64_t orc_enumerate_hands(int32_t ncards, uint64_t first, uint64_t count, uint8_t* out);
int32_t orc_combo_index(int32_t a, int32_t b);
void orc_combo_cards(int32_t idx, int32_t* hi, int32_t* lo);
int32_t orc_combo_class(int32_t idx);
int32_t orc_parse_range(const char* s, float* out1326);
int32_t orc_parse_cards(const char* s, int32_t* out, int32_t cap);
/* all 7-card hands: counts[9] per category, distinct values written to *distinct */
void orc_count_categories(int32_t ncards, uint64_t counts[9], uint64_t* distinct);
int32_t orc_postflop_actions(int32_t street, int64_t pot, int64_t to_call, int64_t current_bet, int64_t max_to,
                             int64_t min_raise_to, uint32_t num_raises, const orc_menu_c* menu, int32_t* kinds,
                             int64_t* tos, int32_t cap);
void orc_discount(uint32_t t, double alpha, int32_t beta_some, double beta, double gamma, int32_t restart,
                  int32_t unweighted, float out[3]);
uint32_t orc_next_check_gap(int32_t has_prev, uint32_t t0, double e0, uint32_t t, double e, double target, uint32_t max_gap);

/* solver */
void* orc_solver_new(const orc_config_c* cfg, char* err, int32_t errlen);
int32_t orc_estimate(const orc_config_c* cfg, orc_stats_c* out, char* err, int32_t errlen);
void orc_solver_free(void* s);
void orc_solver_set_parallel(void* s, int32_t on, int32_t threads);
void orc_solver_set_batch_river(void* s, int32_t on);
void orc_solver_stats(void* s, orc_stats_c* out);
uint64_t orc_solver_storage_len(void* s);
uint64_t orc_solver_num_nodes(void* s);
int32_t orc_solver_num_hands(void* s, int32_t p);
double orc_solver_num_combinations(void* s);
void orc_solver_hands(void* s, int32_t p, uint16_t* combos, float* weights);
void orc_solver_get_storage(void* s, float* regrets, float* sums);
int32_t orc_solver_set_storage(void* s, const float* regrets, const float* sums); /* -1: 16-bit storage */
int32_t orc_solver_is_compressed(void* s);
/* rand 0.9.5's StdRng pieces and sampling.rs's card sample (test hooks) */
void orc_chacha_block(int32_t rounds, const uint32_t* in, uint32_t* out);
void orc_stdrng_u64(uint64_t seed, uint64_t n, uint64_t* out);
/* PCG32 words as rand_core's seed_from_u64 draws them, from a state and increment. */
void orc_pcg32_words(uint64_t state, uint64_t inc, uint64_t n, uint32_t* out);
/* StdRng's stream from a key and round count: n next_u32 words, n / 2 next_u64 draws. */
void orc_stream_words(const uint32_t* key, int32_t rounds, uint64_t n, uint32_t* out, uint64_t* out64);
int32_t orc_sample_cards(const int32_t* available, int32_t n, int32_t k, uint64_t seed, int32_t* out);
/* one block through the 16-bit codec from zero storage (rq: regrets, sq: sums) */
void orc_encode_block(int32_t sums, const float* v, uint64_t n, int16_t* rq, uint16_t* sq, float* scale);
int32_t orc_solver_get_storage_q(void* s, int16_t* regrets, uint16_t* sums, float* regret_scale, float* strat_scale);
int32_t orc_solver_set_storage_q(void* s, const int16_t* regrets, const uint16_t* sums, const float* regret_scale,
                                 const float* strat_scale);
uint32_t orc_solver_iterations(void* s);
void orc_solver_set_iterations(void* s, uint32_t t);
/* runs `count` DCFR iterations (no exploitability checks) */
void orc_solver_iterate(void* s, uint32_t count);
double orc_solver_exploitability(void* s);
double orc_solver_exploitability_f64(void* s);
/* mode 0 cfr (iteration t discount), 1 best response, 2 average */
void orc_solver_root_values(void* s, int32_t mode, int32_t f64, double out[2]);
/* values of player p at the root under `mode` (CFR mode updates storage) */
void orc_solver_traverse_root(void* s, int32_t p, int32_t mode, uint32_t t, float* out);
void orc_solver_traverse_root_f64(void* s, int32_t p, int32_t mode, double* out);
void orc_solver_solve(void* s, uint32_t max_iterations, double target_pct, uint32_t check_every, orc_progress_c* out,
                      orc_progress_c* history, int32_t history_cap, int32_t* history_len);
/* tree export: per node kind (0 p0,1 p1,2 chance,3 fold0,4 fold1,5 showdown), street, num_children,
 * first_child, invest[2], offset, slot, rake, chance_factor, edge card */
void orc_solver_nodes(void* s, uint8_t* kind, uint8_t* street, uint16_t* num_children, uint32_t* first_child,
                      int32_t* invest, uint64_t* offset, uint32_t* slot, float* rake, float* chance_factor,
                      int32_t* edge_card, int32_t* edge_act_kind, int64_t* edge_to);
int32_t orc_solver_num_boards(void* s);
void orc_solver_board(void* s, int32_t i, int32_t* cards5);
/* query: fills counts; arrays sized by caller (na*n). returns 0 ok, else writes err */
int32_t orc_solver_query(void* s, const orc_step_c* path, int32_t npath, int32_t* player, int32_t* na, int32_t* n,
                         int32_t* act_kinds, int64_t* act_tos, uint16_t* combos, float* strategy, float* ev, float* reach,
                         double reach_fraction[2], int64_t* pot, int64_t invest[2], int32_t cap_actions, int32_t cap_hands,
                         char* err, int32_t errlen);
int32_t orc_solver_ranges_after(void* s, const orc_step_c* path, int32_t npath, float* out0, float* out1, char* err,
                                int32_t errlen);

/* terminal evaluators on an explicit hand set (for brute-force pins) */
void* orc_handset_new(const float* r0, const float* r1, const int32_t* board, int32_t nboard);
void orc_handset_free(void* h);
int32_t orc_handset_len(void* h, int32_t p);
void orc_handset_cards(void* h, int32_t p, uint8_t* cards2);
double orc_handset_num_combinations(void* h);
void orc_handset_compatible_reach(void* h, int32_t p, const float* opp, double* out);
void orc_eval_fold(void* h, int32_t p, const float* opp, double payoff, float* out);
/* board is the final 5-card board of the showdown */
void orc_eval_showdown(void* h, const int32_t* board5, int32_t p, const float* opp, double win, double lose, double tie,
                       float* out);
void orc_fold_batch1(void* h, int32_t p, const float* opp, double payoff, float* out);
void orc_showdown_batch1(void* h, const int32_t* board5, int32_t p, const float* opp, double win, double lose, double tie,
                         int32_t has_fold, double fold_pay, float* out, float* fold_out);

#ifdef __cplusplus
}
#endif
#endif
