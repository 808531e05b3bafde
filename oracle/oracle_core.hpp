// CPU oracle for the heads-up DCFR postflop solve.
//
// TEST INFRASTRUCTURE ONLY. This is a C++ restatement of the reference's
// CPU solver (AhanDatta/Poker-Trainer, crates/holdem and crates/solver),
// written from the Rust sources read as text. It is the parity checker for
// the HIP engine and the timed CPU baseline. Only tests/, the smoke entry
// point and the benchmark's cpu_baseline leg may load it; the product
// library never links or calls it.
//
// Every function follows the reference operation for operation (same f32/f64
// types, same summation order, no FMA contraction: build with
// -ffp-contract=off) so that its results are the reference CPU solve's bits.
//
// Pins (checked by tests/test_oracle_*.py):
//   cards/combos  closed forms, bijection, class indices (combo.rs:213-248)
//   evaluator     5-card category counts (eval.rs:245-265), 7-card category
//                 totals over all 133,784,560 hands, distinct value counts,
//                 best-5-of-7 brute force (eval.rs:267-288)
//   menu          the reference's menu cases (menu.rs:412-493)
//   ranges        notation counts (range.rs:275-300)
//   card removal  O(n^2) brute force for compat/fold/showdown (proptests.rs:94-184)
//   DCFR factors  closed forms (cfr.rs:83-107)
//   solver        toy game, all-in equity, ties, rake, convergence
//                 (tests.rs:42-149,305-345), zero-sum / exploitability >= 0
//                 invariants (proptests.rs:224-259), batched == recursive
//                 (proptests.rs:266-302), check schedule (tests.rs:347-359)
// Parity unpinned (no artifact can pin it without running the Rust build):
//   the exact bit patterns of multi-iteration solves beyond the invariants
//   above, and suit isomorphism (iso.rs) beyond the lossless check.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

namespace orc {

using Chips = int64_t;

// ---------------------------------------------------------------- cards
// crates/holdem/src/card.rs:25-89: card = rank*4 + suit, rank 0=2..12=A,
// suit 0=c,1=d,2=h,3=s.
inline int rank_of(int c) { return c >> 2; }
inline int suit_of(int c) { return c & 3; }
inline uint64_t card_bit(int c) { return 1ull << c; }

int parse_card(const std::string& s);              // card.rs:113-126
std::vector<int> parse_cards(const std::string& s); // card.rs:129-136

// crates/holdem/src/combo.rs:12-17
inline int combo_index(int a, int b) {
  int lo = a < b ? a : b, hi = a < b ? b : a;
  return hi * (hi - 1) / 2 + lo;
}
// combo.rs:19-41: (higher-index card, lower-index card)
struct ComboCards { uint8_t hi, lo; };
const ComboCards& combo_cards(int idx);
int combo_class(int idx);  // combo.rs:464-485 (grid index)

struct HandClass {  // combo.rs:56-207
  int hi, lo; bool suited;
  static HandClass make(int r1, int r2, bool suited);
  static HandClass from_grid(int g);
  int grid() const;
  std::vector<int> combos() const;
};

// ---------------------------------------------------------------- evaluator
// crates/holdem/src/eval.rs:70-198
uint64_t eval_mask_bit(int card);
uint32_t evaluate_mask(uint64_t mask);
uint32_t evaluate_cards(const int* cards, int n);

// ---------------------------------------------------------------- ranges
// crates/holdem/src/range.rs:201-264; returns false on a parse error.
bool parse_range(const std::string& s, std::vector<float>& out);

// ---------------------------------------------------------------- menu
enum Street { Preflop = 0, Flop = 1, Turn = 2, River = 3 };
inline int board_len(Street s) { return s == Preflop ? 0 : (int)s + 2; }
inline Street street_from_board(int n) { return n == 0 ? Preflop : n == 3 ? Flop : n == 4 ? Turn : River; }

enum ActKind { Fold = 0, Check = 1, Call = 2, Raise = 3 };
struct Action {
  ActKind kind; Chips to;
  bool operator==(const Action& o) const { return kind == o.kind && (kind != Raise || to == o.to); }
};

struct BetSize { enum Kind { Pot = 0, Mult = 1, AllIn = 2 } kind; double v; };
bool parse_bet_size(const std::string& s, BetSize& out);  // menu.rs:26-51

struct StreetSizes { std::vector<BetSize> bet, raise; };
struct PostflopMenu {  // menu.rs:100-133
  StreetSizes flop, turn, river;
  uint32_t max_raises = 3;
  double add_allin_threshold = 1.5, force_allin_threshold = 0.15, merge_threshold = 0.1;
  static PostflopMenu defaults();
  const StreetSizes& street(Street s) const { return s == Turn ? turn : s == River ? river : flop; }
};

struct Situation {  // menu.rs:191-217
  Street street; Chips pot, to_call, current_bet, max_to, min_raise_to;
  uint32_t num_raises, limpers; Chips big_blind;
  bool can_raise() const { return max_to > current_bet; }
};
std::vector<Action> postflop_actions(const Situation& sit, const PostflopMenu& menu);  // menu.rs:231-308

// ---------------------------------------------------------------- betting
struct BettingState {  // crates/solver/src/tree.rs:17-126
  Street street; Chips pot; Chips commit[2]; Chips stack[2];
  Chips current_bet, last_raise_inc; uint32_t num_raises; int to_act; bool needs[2];
  static BettingState street_start(Street s, Chips pot, Chips s0, Chips s1, Chips bb);
  Situation situation(Chips bb) const;
};
enum TransKind { Continue = 0, FoldT = 1, StreetOver = 2 };
struct Transition { TransKind kind; int folder; BettingState st; Chips put; };
Transition apply(const BettingState& s, const Action& a);

// ---------------------------------------------------------------- tree
enum Kind : uint8_t { KPlayer0 = 0, KPlayer1 = 1, KChance = 2, KFold0 = 3, KFold1 = 4, KShowdown = 5 };
inline bool is_player(Kind k) { return k <= KPlayer1; }
inline bool is_fold(Kind k) { return k == KFold0 || k == KFold1; }

struct Node {  // tree.rs:154-205
  Kind kind; Street street; uint16_t num_children; uint32_t first_child;
  int32_t invest[2]; uint64_t offset; uint32_t slot; float rake; float chance_factor;
  // edge: 0 root, 1 action, 2 card
  uint8_t edge_kind; Action edge_action; int edge_card;
};

struct TreeStats {  // tree.rs:247-274
  uint64_t player_nodes = 0, chance_nodes = 0, terminal_nodes = 0, storage_cells = 0;
  bool compressed = false;
  uint64_t total_nodes() const { return player_nodes + chance_nodes + terminal_nodes; }
  // size_of::<Node>() of the Rust arena node (tree.rs:264-268): 56 bytes.
  uint64_t memory_bytes() const { return storage_cells * (compressed ? 4 : 8) + total_nodes() * 56; }
  uint64_t memory_bytes_f32() const { return storage_cells * 8 + total_nodes() * 56; }
};

struct ChanceInfo { std::vector<std::pair<uint16_t, uint8_t>> iso; std::vector<int> iso_cards; };

struct Tree {
  std::vector<Node> nodes;
  std::vector<ChanceInfo> chance;
  std::vector<std::vector<int>> boards;
  uint64_t storage_len = 0;
  uint32_t num_player_nodes = 0;
  std::vector<Action> actions(size_t id) const;
};

// ---------------------------------------------------------------- hands
constexpr uint32_t NONE = 0xFFFFFFFFu;
struct HandSet {  // crates/solver/src/hands.rs:14-126
  std::vector<uint16_t> combos[2];
  std::vector<std::pair<uint8_t, uint8_t>> cards[2];
  std::vector<float> weights[2];
  std::vector<uint32_t> same[2];
  std::vector<std::vector<uint32_t>> blocked[2];  // [card] -> hands
  std::vector<uint32_t> lookup[2];
  double num_combinations = 0;
  size_t len(int p) const { return combos[p].size(); }
  std::vector<double> compatible_reach(int p, const float* opp_reach) const;
};
HandSet make_handset(const std::vector<float>* ranges, uint64_t board, float prune,
                     const std::vector<int>* keep);

struct Entry {  // crates/solver/src/eval.rs:11-18, ordered (strength, idx, c1, c2)
  uint32_t strength; uint16_t idx; uint8_t c1, c2;
  bool operator<(const Entry& o) const {
    if (strength != o.strength) return strength < o.strength;
    if (idx != o.idx) return idx < o.idx;
    if (c1 != o.c1) return c1 < o.c1;
    return c2 < o.c2;
  }
};
struct StrengthTable { std::vector<Entry> lists[2]; };
StrengthTable make_table(const HandSet& hs, const std::vector<int>& board);

// ---------------------------------------------------------------- iso
struct Isomorphism {  // crates/solver/src/iso.rs:236-301
  bool enabled = false;
  bool range_symmetric[6] = {false, false, false, false, false, false};
  std::vector<uint16_t> perm[2][6];
  void chance_cards(const std::vector<int>& board, const std::vector<int>& available,
                    std::vector<int>& canon, std::vector<std::array<int, 3>>& refs) const;
};
Isomorphism make_iso(const HandSet& hs, bool enabled);
extern const int PAIRS[6][2];

// ---------------------------------------------------------------- dcfr
struct DcfrParams { double alpha = 1.5; bool beta_some = true; double beta = 0.0; double gamma = 3.0;
                    bool restart_average = true; bool unweighted_average = true; };
struct Discount { float pos, neg, strat; bool reach_weighted; };
Discount make_discount(uint32_t t, const DcfrParams& p);  // cfr.rs:83-107

enum ModeKind { MCfr = 0, MBest = 1, MAvg = 2 };
struct Mode { ModeKind kind; Discount d; };

// ---------------------------------------------------------------- solver
enum class Backend { Cpu = 0, Gpu = 1, Auto = 2 };

struct SolveConfig {  // crates/solver/src/lib.rs:73-135
  std::vector<int> board;
  std::vector<float> ranges[2];
  BettingState root;
  PostflopMenu menu;
  bool has_future_menu = false;
  PostflopMenu future_menu;
  Chips big_blind = 100;
  float prune = 0.0f;
  std::vector<int> keep[2];
  uint64_t max_memory_bytes = 8ull << 30;
  double rake_rate = 0.0, rake_cap = 0.0;
  int storage_mode = 0;  // storage.rs:74-81: 0 F32, 1 Compressed, 2 Auto (compressed above 1 GB of f32)
  bool isomorphism = true;
  // tree.rs:128-136: 0 Exact, 1 Sampled { turn, river, seed }
  int chance_mode = 0;
  uint32_t sample_turn = 0, sample_river = 0;
  uint64_t sample_seed = 0;
  DcfrParams dcfr;
  Backend backend = Backend::Cpu;
};

struct SolveOptions { uint32_t max_iterations = 1000; double target_exploitability_pct = 0.5; uint32_t check_every = 10; };
struct Progress { uint32_t iteration, max_iterations; double exploitability, exploitability_pct; bool finished; };

uint32_t next_check_gap(bool has_prev, uint32_t t0, double e0, uint32_t t, double e, double target, uint32_t max_gap);

struct NodeInfo {  // lib.rs:219-275
  int player;
  std::vector<Action> actions;
  std::vector<uint16_t> combos;
  std::vector<float> strategy, ev, reach;
  double reach_fraction[2];
  Chips pot; Chips invest[2];
};

enum StepKind { StepAction = 0, StepCard = 1 };
struct Step { StepKind kind; Action action; int card; };

// 16-bit storage (storage.rs:87-106): regrets as i16 and strategy sums as
// u16, each player node's block with one f32 scale (its largest |value|);
// a stored q stands for q * (scale / MAX).
struct Storage16 {
  std::vector<int16_t> regrets;
  std::vector<uint16_t> strat;
  std::vector<float> regret_scale, strat_scale;  // per player node (Node::slot)
};

struct Solver;
struct Ctx {
  const Tree* tree; std::vector<float>* regrets; std::vector<float>* strat;
  Storage16* q;  // null: f32 storage in regrets / strat
  const HandSet* hands; const std::vector<StrengthTable>* tables; const Isomorphism* iso;
  double half_pot, norm; bool batch_river; bool parallel;
};

struct Solver {
  SolveConfig config;
  HandSet hands;
  Isomorphism iso;
  Tree tree;
  std::vector<float> regrets, strat;  // f32 storage (empty when compressed)
  Storage16 q;                        // 16-bit storage (empty when f32)
  std::vector<StrengthTable> tables;
  TreeStats stats;
  uint32_t iterations = 0;
  double exploitability_ = std::numeric_limits<double>::infinity();
  bool batch_river = true;
  bool parallel = true;

  Ctx ctx();
  bool compressed() const { return stats.compressed; }
  bool raked() const { return config.rake_rate > 0.0 && config.rake_cap > 0.0; }
  void cpu_iteration();
  double exploitability();
  double exploitability_f64();
  void root_values(ModeKind m, double out[2]);
  void root_values_f64(ModeKind m, double out[2]);
  Progress solve(const SolveOptions& opts, std::vector<Progress>* history);
  std::vector<float> average_strategy(size_t id, size_t n);
  // query / ranges_after; return "" on success or an error message
  std::string walk(const std::vector<Step>& path, size_t& node, std::vector<float> reach[2],
                   std::vector<uint32_t> to_canon[2]);
  std::string query(const std::vector<Step>& path, NodeInfo& out);
  std::string ranges_after(const std::vector<Step>& path, std::vector<float> out[2]);
};

// rand 0.9.5's StdRng (ChaCha12Rng seeded by seed_from_u64) and the
// reference's stratified card sample (sampling.rs)
void chacha_block(int rounds, const uint32_t in[16], uint32_t out[16]);
// one PCG32 word as rand_core's seed_from_u64 draws it: the state advanced
// (state * 6364136223846793005 + inc), then the XSH RR output of the new state
uint32_t pcg32_next(uint64_t& state, uint64_t inc);
struct StdRng {
  uint32_t key[8];
  uint64_t counter = 0;
  uint32_t buf[64];
  size_t idx = 64;
  int rounds = 12;
  explicit StdRng(uint64_t seed);
  // the same stream from a given key and round count (ChaCha20 for the
  // published keystream vectors)
  StdRng(const uint32_t k[8], int rounds);
  void refill();
  uint32_t next_u32();
  uint64_t next_u64();
  double next_f64();
};
std::vector<int> sample_cards(const std::vector<int>& available, size_t k, StdRng& rng);

// One block through the 16-bit codec from zero storage: the stored integers
// (rq for regrets, sq for sums) and the block's new scale.
void encode_block(bool sums, const float* v, size_t n, int16_t* rq, uint16_t* sq, float* scale);

// Returns "" on success, else an error (SolverError Display text).
std::string estimate(const SolveConfig& cfg, TreeStats& out);
std::string make_solver(const SolveConfig& cfg, Solver& out);

// traversal entry (cfr.rs:258-288), f32 values
void traverse(const Ctx& ctx, size_t id, int p, const float* opp_reach, size_t nq, Mode mode, float* out, size_t n);
// the same traversal with FP64 reach and values (BestResponse / Average only)
void traverse_f64(const Ctx& ctx, size_t id, int p, const double* opp_reach, size_t nq, ModeKind mode, double* out, size_t n);

// terminal evaluators (crates/solver/src/eval.rs)
void eval_fold(const HandSet& hs, int p, const float* opp_reach, double payoff, float* out);
void eval_showdown(const HandSet& hs, const StrengthTable& t, int p, const float* opp_reach,
                   double win, double lose, double tie, float* out);
void fold_batch(const HandSet& hs, int p, int k, const float* const* reaches, const double* payoffs, float* const* outs);
void showdown_batch(const HandSet& hs, const StrengthTable& t, int p, int k, const float* const* reaches,
                    const double* win, const double* lose, const double* tie, float* const* outs,
                    const int* has_fold, const double* fold_pay, float* const* fold_outs);

}  // namespace orc
