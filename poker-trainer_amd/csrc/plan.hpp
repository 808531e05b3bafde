// The GPU plan: the game-tree arena flattened into structure-of-arrays that
// the HIP kernels walk level by level.
//
// Segments. The tree is cut at chance nodes: a segment is the tree root or a
// chance node's child together with every descendant reachable without
// crossing another chance node (chance nodes are the segment's leaves; their
// children root deeper segments). Segments are grouped by depth (the number
// of chance nodes above them). One traversal of player p runs
//   down(depth 0..D-1) -> terminals -> up(depth D-1..0)
// where a down kernel gives one workgroup per segment (the opponent's reach
// pushed down a level of the opponent's nodes at a time, with whole-row
// "alive" tests as block reductions), the terminal kernel gives one workgroup per terminal,
// and an up kernel gives one workgroup per (segment, 64-hand tile) (values
// pulled up in reverse member order, regret update fused).
//
// Reach slots (per traverser p, opponent o = p^1): slot 0 is the root; every
// child of an o-acting node and every chance child gets a new slot; children
// of p-acting nodes share their parent's slot (crates/solver/src/river.rs:
// 92-161 does the same inside river subtrees). Value rows exist for
// terminals and segment roots only; everything else lives in LDS.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "host.hpp"

namespace pts {

// Terminal flags
constexpr uint8_t kTermBatched = 1;   // inside a river subtree (river.rs batched formulas)
constexpr uint8_t kTermPaired0 = 2;   // fold read off the sibling showdown's sums (traverser 0)
constexpr uint8_t kTermPaired1 = 4;   // (traverser 1)
// Decision-node flag: its regrets are pinned (the up pass computes its value
// and leaves its regrets as they are)
constexpr uint8_t kNodeFrozen = 8;

// One member of a segment, everything the down / up kernels read about it
// (staged in LDS once per workgroup). Children of a decision node are
// consecutive members; children of a chance node are consecutive segment
// roots, with consecutive reach slots and value rows.
struct MemRec {
  uint64_t offset;     // decision: storage offset; chance: first child's node id
  uint32_t node;       // arena id
  uint32_t vslot;      // value row (terminal / segment root); chance: first child's row
  uint32_t slot[2];    // reach slot, per traverser
  uint32_t cslot[2];   // first child's reach slot, per traverser (fresh slots only)
  uint16_t clocal;     // local index of the first child (decision nodes)
  uint16_t src[2];     // member whose slot decides this node's alive flag, per traverser
  uint8_t kind, na;
  int8_t card;         // chance child: the dealt card
  uint8_t tflags;
  uint16_t spot;       // the spot (tree) of a batch it belongs to
  uint32_t pslot;      // decision: its index among the player nodes (16-bit storage scale slot)
};
static_assert(sizeof(MemRec) == 48, "MemRec layout");

struct TermRec {
  uint32_t node, slot[2], vslot, board;
  int32_t invest[2];
  float rake;
  uint8_t kind, tflags;
  uint16_t spot;
  uint8_t pad[4];
};
static_assert(sizeof(TermRec) == 40, "TermRec layout");

// A run of consecutive terminals that read the same strength tables (one
// river board) or none (folds before the river, summed in hand order): the
// fast terminal kernel loads the tables once per run.
struct TermBatch {
  uint32_t t0;     // first terminal record
  uint32_t cnt;    // terminals in the run
  uint32_t board;  // strength-table board, or kNone for hand order
  uint32_t spot;
};
static_assert(sizeof(TermBatch) == 16, "TermBatch layout");
// terminals per run: long runs load a river board's tables and first reach
// row once for all of the board's terminals in a chunk (the flop's terminal
// pass: 64.0 ms per iteration at 16, 59.9 at 64, 58.9 at 256)
constexpr uint32_t kBatchMax = 256;

// A leaf of a depth-limited tree: valued by the value network per dealt card.
struct LeafRec {
  uint32_t node, vslot, slot[2];
  uint32_t pot, stack;  // chips in the pot, the smaller stack behind
  float factor;         // the cut chance node's reach factor
  uint16_t local, spot; // member index in its segment; spot
};
static_assert(sizeof(LeafRec) == 32, "LeafRec layout");

struct PlanInput {
  size_t num_nodes = 0;
  const uint8_t* kind = nullptr;
  const uint8_t* street = nullptr;
  const uint16_t* num_children = nullptr;
  const uint32_t* first_child = nullptr;
  const int32_t* invest = nullptr;     // 2 per node
  const uint64_t* offset = nullptr;    // player: storage offset; showdown: board index
  const float* rake = nullptr;
  const float* chance_factor = nullptr;
  const int8_t* card = nullptr;        // chance child: dealt card, else -1
  uint32_t num_boards = 0;
  const int32_t* boards = nullptr;     // 5 per board
  uint32_t num_hands[2] = {0, 0};
  const uint8_t* hand_cards[2] = {nullptr, nullptr};
  const float* hand_weights[2] = {nullptr, nullptr};
  const uint32_t* hand_same[2] = {nullptr, nullptr};
  double half_pot = 0, num_combinations = 0;
  uint64_t storage_len = 0;
  bool batch_river = true;
  int64_t root_stack[2] = {0, 0};  // chips behind at the root (for leaves)
  // suit isomorphism (none when iso_refs is null): per node its merged cards
  // (chance nodes), and per player 6 hand permutations of num_hands each
  const uint32_t* iso_first = nullptr;
  const uint16_t* iso_cnt = nullptr;
  const IsoRef* iso_refs = nullptr;
  const uint16_t* iso_perm[2] = {nullptr, nullptr};
  // per node its player-node index (player nodes; null: no 16-bit storage)
  const uint32_t* pslot = nullptr;
  uint32_t num_pslots = 0;
};

struct Plan {
  size_t num_nodes = 0;
  uint32_t n[2] = {0, 0};
  // per node
  std::vector<uint8_t> kind, na, tflags;
  std::vector<uint32_t> first_child, vslot, tboard;
  std::vector<uint32_t> subtree_end;  // one past the largest node id in each node's subtree
  std::vector<uint32_t> seg_of;       // segment index of every node
  std::vector<uint64_t> offset;
  std::vector<int32_t> invest;
  std::vector<float> rake, factor;
  std::vector<int8_t> card;
  std::vector<uint32_t> rslot;       // [2][num_nodes]
  std::vector<uint16_t> local;       // index in the segment's member list
  std::vector<uint16_t> alive_src;   // [2][num_nodes]: member whose slot decides this node's alive flag
  // segments, ordered by depth
  std::vector<uint32_t> seg_root, seg_parent, seg_mem_off, seg_mem_cnt;
  std::vector<uint32_t> depth_off;   // segments of depth d: [depth_off[d], depth_off[d+1])
  std::vector<uint32_t> members;
  std::vector<uint32_t> terminals;
  std::vector<MemRec> mrec;          // parallel to members
  std::vector<TermRec> trec;         // parallel to terminals
  std::vector<uint32_t> seg_pslot;   // [2][segments]: the parent chance node's reach slot
  std::vector<float> seg_factor;     // the parent chance node's factor
  // chance nodes grouped by the depth of their segment: their values are
  // summed by a wide kernel before that depth's up pass
  std::vector<uint32_t> chance_off;  // [depth + 1]
  std::vector<MemRec> crec;          // one per chance node (vslot: its own row; offset: first child; cslot: children)
  std::vector<uint32_t> chance_child_vslot;  // first child's value row, parallel to crec
  // the cards each chance record's node merges by suit isomorphism:
  // ciso[ciso_off[i], ciso_off[i + 1]) (ciso_off has crec.size() + 1 entries)
  std::vector<uint32_t> ciso_off;
  std::vector<IsoRef> ciso;
  std::vector<uint16_t> iso_perm[2];  // 6 x n[p] hand permutations (empty without isomorphism)
  uint32_t num_slots[2] = {0, 0}, num_vslots = 0, max_members = 0;
  // chunks: the subtrees below the root street's chance nodes (the depth-1
  // segments and their descendants), run in groups that share scratch.
  // Reach slots below slot_s0[p] and value rows below val_v0 are resident;
  // chunk c owns slots [slot_coff[p][c], slot_coff[p][c+1]) and rows
  // [val_coff[c], val_coff[c+1]).
  uint32_t num_chunks = 0;
  std::vector<uint32_t> chunk_of;      // per node: chunk, or kNone in the top segment
  std::vector<uint32_t> seg_coff;      // [d * (C+1) + c], d >= 1: first segment of chunk c at depth d
  std::vector<uint32_t> ch_coff;       // [d * (C+1) + c], d >= 1: first chance record of chunk c at depth d
  std::vector<uint32_t> term_coff;     // [c]: first terminal of chunk c ([0]: after the top segment's)
  std::vector<uint32_t> slot_coff[2];  // [c], c <= C
  std::vector<uint32_t> val_coff;      // [c], c <= C
  uint32_t slot_s0[2] = {0, 0}, val_v0 = 0;
  // terminal runs: the top segment's first, then chunk by chunk
  // ([tb_coff[c], tb_coff[c+1]) for chunk c; tb_coff[0] = the top's count)
  std::vector<TermBatch> tbatch;
  std::vector<uint32_t> tb_coff;
  std::vector<LeafRec> leaves;        // depth-limited trees: the root street's cut points
  // spots: one for a single tree; a batch (merge_plans) keeps every spot's
  // own plan in `parts` for its hands and strength tables
  uint32_t num_spots = 1;
  std::vector<Plan> parts;
  std::vector<uint32_t> spot_root_node, spot_root_vslot, spot_root_slot[2];
  const Plan& spot(uint32_t k) const { return parts.empty() ? *this : parts[k]; }
  // hands
  std::vector<uint8_t> cards[2];
  std::vector<float> weights[2];
  std::vector<uint32_t> same[2];
  std::vector<uint32_t> lookup[2];   // combo (hi * (hi - 1) / 2 + lo) -> hand, or kNone
  std::vector<uint32_t> bl_off[2];   // 53 per player: hands holding each card (ascending)
  std::vector<uint16_t> bl_idx[2];
  // strength tables, per board b and player q (the player whose hands are listed)
  uint32_t num_boards = 0;
  std::vector<uint32_t> ent_off[2], ent_cnt[2];  // [b]
  std::vector<uint16_t> ent_idx[2];              // hand index per entry, ascending (strength, index)
  std::vector<uint32_t> ent_str[2];              // strength per entry
  std::vector<uint16_t> cl_off[2];               // [b*53 + c]: start of card c's list in the board block
  std::vector<uint16_t> cl_idx[2];               // hand index, ascending strength, per card (block of 2*cnt)
  std::vector<uint32_t> cl_base[2];              // [b]: block start in cl_idx
  // per board b, for q as "mine" against the other player's list: per entry
  // lt/le counts and per-card lt/le counts (u16 x 6), and the entry of each hand
  std::vector<uint16_t> my_cnt[2];   // 6 per entry, at ent_off
  std::vector<uint16_t> hand_ent[2]; // [b * n_q + hand] -> entry or 0xFFFF
  // the same counts keyed by hand: [(b * n_q + hand) * 8 + k], k = lt, le,
  // lt(c1), le(c1), lt(c2), le(c2), then two zero pads; lt = 0xFFFF for
  // hands blocked by the board
  std::vector<uint16_t> hand_cnt[2];
  uint32_t max_entries = 0;
  double half_pot = 0, norm = 0;
  uint64_t storage_len = 0;
  uint32_t num_pslots = 0;  // player nodes with a scale slot (0: no 16-bit storage)
  bool batch_river = true;
  uint32_t depths() const { return (uint32_t)depth_off.size() - 1; }
};

// Builds and validates the plan; returns "" or the reason it is invalid.
std::string build_plan(const PlanInput& in, Plan& out);

// Terminal runs of at most `bmax` terminals (rebuilds tbatch / tb_coff), and
// the default length for a tree of this size.
void build_runs(Plan& P, uint32_t bmax);
uint32_t default_run_length(const Plan& P);

// A batch of independent spots as one plan: segments of every spot merged by
// depth, ids shifted, storage concatenated in spot order, every row
// resident (no chunk groups). The spots' own plans move into `parts`.
std::string merge_plans(std::vector<Plan>&& parts, Plan& out);

// Chunk groups [c0, c1) and the scratch rows they need.
struct ChunkGroup {
  uint32_t c0, c1;
};
struct ScratchRows {
  uint64_t reach_bytes = 0, val_bytes = 0, alive_rows = 0;
  uint64_t slot_rows[2] = {0, 0}, val_rows = 0;  // rows per traverser / value rows
};
// Consecutive chunks whose own reach slots and value rows (of `es`-byte
// elements, rows of `rowlen`) fit `budget` beside the resident ones; a chunk
// that alone exceeds it gets a group of its own.
ScratchRows partition_chunks(const Plan& P, size_t es, uint32_t rowlen, uint64_t budget, std::vector<ChunkGroup>& groups);
// Checks that every slot and row a grouped traversal touches maps inside the
// scratch rows (the shift the kernels apply); returns "" or the first failure.
std::string check_groups(const Plan& P, const std::vector<ChunkGroup>& groups, const ScratchRows& rows);

}  // namespace pts
