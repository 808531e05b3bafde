// Host side of the MI355X heads-up DCFR solver: cards, hand evaluator, bet
// menu, the game-tree arena, hand sets and strength tables.
//
// The arena (node order, storage offsets, chance factors) is the one the
// reference solver builds (crates/solver/src/tree.rs:403-574), because the
// regret / strategy-sum layout `offset + action * hands + hand` is the data
// format the engine shares with its caller (crates/solver/src/storage.rs:87-106).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace pts {

using Chips = int64_t;
constexpr uint32_t kNone = 0xFFFFFFFFu;

enum Street : uint8_t { kPreflop = 0, kFlop = 1, kTurn = 2, kRiver = 3 };

// ----------------------------------------------------------------- cards
// Card = rank*4 + suit (crates/holdem/src/card.rs:25-69). Combo index of
// {a, b} = hi*(hi-1)/2 + lo (crates/holdem/src/combo.rs:12-17).
inline int combo_of(int a, int b) { int lo = a < b ? a : b, hi = a ^ b ^ lo; return hi * (hi - 1) / 2 + lo; }
void combo_pair(int combo, int& hi, int& lo);

// 7-card strength, category << 26 | payload (crates/holdem/src/eval.rs:142-198)
uint32_t hand_strength(uint64_t lane_mask);
inline uint64_t lane_bit(int card) { return 1ull << ((card & 3) * 16 + (card >> 2)); }

// ----------------------------------------------------------------- menu
struct Size { int kind; double v; };  // 0 pot fraction, 1 multiple, 2 all-in
struct Menu {
  std::vector<Size> bet[3], raise[3];  // flop, turn, river
  uint32_t max_raises = 3;
  double add_allin = 0, force_allin = 0, merge = 0;
};
enum Act : uint8_t { kFold = 0, kCheck = 1, kCall = 2, kRaise = 3 };
struct Action { Act kind; Chips to; };

struct Betting {  // crates/solver/src/tree.rs:17-126
  Street street; Chips pot; Chips commit[2]; Chips stack[2];
  Chips current_bet, last_raise_inc; uint32_t num_raises; int to_act; bool needs[2];
};
// Actions at a betting state (crates/holdem/src/menu.rs:264-392 via tree.rs:298-300)
std::vector<Action> menu_actions(const Betting& b, const Menu& m, Chips big_blind);
// 0 continue, 1 fold, 2 street over; `put` = chips the actor adds
int betting_apply(const Betting& b, const Action& a, Betting& next, Chips& put);

// ----------------------------------------------------------------- tree
// kLeaf: the end of the root street's betting in a depth-limited tree (the
// next street's cards and play are valued by the value network)
enum Kind : uint8_t { kP0 = 0, kP1 = 1, kChance = 2, kFold0 = 3, kFold1 = 4, kShowdown = 5, kLeaf = 6 };
inline bool is_player(uint8_t k) { return k <= kP1; }
inline bool is_terminal(uint8_t k) { return k >= kFold0; }

struct Node {
  uint8_t kind, street;
  uint16_t num_children;
  uint32_t first_child;
  int32_t invest[2];
  uint64_t offset;  // player: storage offset; showdown: board index; leaf: pot (chips)
  uint32_t pslot;   // player: its index among the player nodes (the 16-bit storage's scale slot)
  float rake, chance_factor;
  int8_t card;      // chance child: the dealt card, else -1
  Action action;    // action child: the action taken
  // chance node: its cards merged by suit isomorphism, Tree::iso_refs[iso_first, + iso_cnt)
  uint32_t iso_first;
  uint16_t iso_cnt;
};

// ----------------------------------------------------------------- isomorphism
// Suit pairs (crates/solver/src/iso.rs:174) under which both ranges are
// unchanged (weights within 1e-4 of the largest, iso.rs:245-273), with the
// hand permutation each one induces, per player.
constexpr int kSuitPairs[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
struct Hands;
struct Iso {
  bool enabled = false;
  bool symmetric[6] = {false, false, false, false, false, false};
  std::vector<uint16_t> perm[2][6];  // hand h -> the hand with the pair's suits swapped
  bool any() const {
    for (bool b : symmetric) if (b) return true;
    return false;
  }
};
Iso make_iso(const Hands& h, bool enabled);
// A dealt card merged onto a canonical child: the child's values, with the
// hands permuted by suit pair `pair`, stand for the card's (iso.rs:277-300).
struct IsoRef {
  uint16_t child;  // canonical child index at the chance node
  uint8_t pair;
  int8_t card;
};

struct TreeSpec {
  Street root_street;
  std::vector<int> board;
  Betting root;
  Menu menu, future_menu;
  Chips big_blind = 100;
  double rake_rate = 0, rake_cap = 0;
  size_t hands[2] = {0, 0};
  // cut the tree where the root street's betting ends with chips behind:
  // such a node becomes a leaf (its chance factor as the chance node's)
  bool depth_limit = false;
  // suit isomorphism: chance nodes keep one child per class of equivalent cards
  const Iso* iso = nullptr;
  // tree.rs:128-136, ChanceMode::Sampled: chance nodes deal a stratified
  // sample of sample_turn / sample_river cards (not in all-in runouts)
  bool sampled = false;
  uint32_t sample_turn = 0, sample_river = 0;
  uint64_t sample_seed = 0;
};

// ----------------------------------------------------------------- sampled chance
// The random stream of the reference's sampled mode: rand 0.9.5's StdRng
// (rand_chacha 0.9.0's ChaCha12Rng) seeded by rand_core 0.9.5's
// seed_from_u64, and its f64 draw. ChaCha with 12 rounds (RFC 8439 layout,
// 64-bit block counter in words 12-13, stream 0 in words 14-15), key words
// from PCG32 (XSH RR) steps of the seed; u64 draws take two consecutive
// words, low first; f64 draws are (u64 >> 11) * 2^-53.
class ChaChaRng {
 public:
  explicit ChaChaRng(uint64_t seed);
  uint64_t next_u64();
  double next_f64() { return (double)(next_u64() >> 11) * 0x1.0p-53; }
  // the block function for `rounds` rounds (20 is RFC 8439's ChaCha20)
  static void block(const uint32_t in[16], uint32_t out[16], int rounds);

 private:
  uint32_t next_word();
  uint32_t key_[8];
  uint64_t blocks_ = 0;  // blocks generated so far (the next block's counter)
  uint32_t words_[16];
  uint32_t pos_ = 16;    // next unread word of words_
};

// sampling.rs:9-56: k of the available cards (ascending), each suit given
// its proportional share (largest remainders first, ties broken by a random
// draw), spread over its ranks from a random offset.
std::vector<int> stratified_cards(const std::vector<int>& available, size_t k, ChaChaRng& rng);

struct TreeCounts { uint64_t player = 0, chance = 0, terminal = 0, cells = 0; };

struct Tree {
  std::vector<Node> nodes;
  std::vector<IsoRef> iso_refs;
  std::vector<std::vector<int>> boards;  // full 5-card board of every showdown table
  uint64_t storage_len = 0;
  uint32_t player_nodes = 0;
};

TreeCounts count_tree(const TreeSpec& spec);
Tree build_tree(const TreeSpec& spec);

// ----------------------------------------------------------------- hands
struct Hands {  // crates/solver/src/hands.rs:14-126
  std::vector<uint16_t> combos[2];
  std::vector<uint8_t> cards[2];     // 2 per hand: higher card index first
  std::vector<float> weights[2];
  std::vector<uint32_t> same[2];     // identical combo in the other list
  std::vector<uint32_t> lookup[2];   // combo -> hand
  double num_combinations = 0;
  size_t n(int p) const { return combos[p].size(); }
};
Hands make_hands(const float* const ranges[2], uint64_t board_bits);
std::vector<double> compatible_reach(const Hands& h, int p, const float* opp);

}  // namespace pts
