// C ABI (include/pts.h): the engine layer and the Solver layer that mirrors
// crates/solver/src/lib.rs on top of it.
#include <cmath>
#include <cstdio>
#include <chrono>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>

#include "../../include/pts.h"
#include "engine.hpp"
#include "host.hpp"
#include "net.hpp"
#include "plan.hpp"

using namespace pts;

struct pts_engine {
  std::unique_ptr<Engine> e;
};

struct pts_solver {
  pts_config_c cfg;
  TreeSpec spec;
  Tree tree;
  Hands hands;
  Iso iso;
  TreeCounts counts;
  std::unique_ptr<Engine> engine;
  Plan host_plan;  // kept by host-only solvers (device -1)
  bool compressed = false;  // 16-bit storage (storage.rs:74-81)
  // One engine call at a time: queries may come from reader threads while
  // (or after) another thread solves (crates/trainer/src/study.rs:995-1051)
  std::recursive_mutex mu;
  uint32_t iterations = 0;
  double exploitability = std::numeric_limits<double>::infinity();
};

static pts_status st(Status s) { return (pts_status)s; }
static pts_status bad(const std::string& m) {
  set_error(m);
  return PTS_INVALID_INPUT;
}

// ------------------------------------------------------------ shared helpers
// crates/solver/src/cfr.rs:83-107
static IterParams discount(uint32_t t, const pts_config_c& c) {
  uint32_t tm = t < 1 ? 1 : t;
  double prev = (double)(tm - 1);
  double pa = std::pow(prev, c.alpha);
  float neg = 0.0f;
  if (c.beta_some) {
    double pb = std::pow(prev, c.beta);
    neg = (float)(pb / (pb + 1.0));
  }
  double since = (double)tm - 1.0;
  if (c.restart_average) {
    uint32_t t0 = tm - 1;
    uint32_t p4 = t0 == 0 ? 0 : (1u << ((31 - __builtin_clz(t0)) & ~1u));
    since = (double)(t0 - p4);
  }
  IterParams d;
  d.pos = (float)(pa / (pa + 1.0));
  d.neg = neg;
  d.strat = (float)std::pow(since / (since + 1.0), c.gamma);
  d.reach_weighted = c.unweighted_average ? 0 : 1;
  return d;
}

// crates/solver/src/lib.rs:158-172
static uint32_t check_gap(bool has_prev, uint32_t t0, double e0, uint32_t t, double e, double target, uint32_t max_gap) {
  uint32_t min_gap = std::max(max_gap / 4, 1u);
  if (!has_prev || !(e < e0 && e > target && t > t0 && target > 0.0)) return max_gap;
  double k = std::log(e0 / e) / std::log((double)t / (double)t0);
  double gap = std::ceil((double)t * std::pow(e / target, 1.0 / k) - (double)t);
  if (!std::isfinite(gap)) return max_gap;
  double g = std::max(gap, 0.0);
  uint32_t gi = g >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)g;
  return std::min(std::max(gi, min_gap), max_gap);
}

// f32 normalization of a node's strategy sums (crates/solver/src/cfr.rs:124-154)
static void normalize(float* s, size_t na, size_t n) {
  std::vector<float> inv(s, s + n);
  for (size_t a = 1; a < na; a++)
    for (size_t h = 0; h < n; h++) inv[h] += s[a * n + h];
  for (size_t h = 0; h < n; h++) {
    float i = 1.0f / inv[h];
    inv[h] = (inv[h] > 0.0f && std::isfinite(i)) ? i : 0.0f;
  }
  for (size_t a = 0; a < na; a++)
    for (size_t h = 0; h < n; h++) s[a * n + h] *= inv[h];
  float uni = 1.0f / (float)na;
  for (size_t h = 0; h < n; h++)
    if (inv[h] == 0.0f)
      for (size_t a = 0; a < na; a++) s[a * n + h] = uni;
}

static Menu menu_of(const pts_menu_c& m) {
  Menu r;
  for (int s = 0; s < 3; s++) {
    for (int i = 0; i < m.bet[s].count && i < 8; i++) r.bet[s].push_back({m.bet[s].kind[i], m.bet[s].value[i]});
    for (int i = 0; i < m.raise[s].count && i < 8; i++) r.raise[s].push_back({m.raise[s].kind[i], m.raise[s].value[i]});
  }
  r.max_raises = m.max_raises;
  r.add_allin = m.add_allin_threshold;
  r.force_allin = m.force_allin_threshold;
  r.merge = m.merge_threshold;
  return r;
}

// lib.rs:307-329 + the parts of Solver::new the GPU build supports
static std::string validate(const pts_config_c& c, TreeSpec& spec, Hands& hands) {
  int n = c.board_len;
  if (n < 3 || n > 5) return "invalid board: " + std::to_string(n) + " cards";
  uint64_t bits = 0;
  for (int i = 0; i < n; i++) {
    if (c.board[i] < 0 || c.board[i] >= 52) return "invalid board: bad card";
    if (bits >> c.board[i] & 1) return "invalid board: duplicate cards";
    bits |= 1ull << c.board[i];
  }
  int street = n == 3 ? kFlop : n == 4 ? kTurn : kRiver;
  if (street != c.street) return "invalid board: board doesn't match the root street";
  if (c.prune != 0.0f) return "invalid input: range pruning is not supported";
  if (c.storage_mode < 0 || c.storage_mode > 2) return "invalid input: storage mode";
  if (c.chance_mode < 0 || c.chance_mode > 1) return "invalid input: chance mode";
  if (c.chance_mode == 1 && (c.sample_turn == 0 || c.sample_river == 0))
    return "invalid input: sampled chance deals at least one turn and one river card";
  spec.sampled = c.chance_mode == 1;
  spec.sample_turn = c.sample_turn;
  spec.sample_river = c.sample_river;
  spec.sample_seed = c.sample_seed;
  const float* r[2] = {c.ranges[0], c.ranges[1]};
  hands = make_hands(r, bits);
  for (int p = 0; p < 2; p++)
    if (hands.n(p) == 0) return "player " + std::to_string(p) + " has an empty range";
  if (hands.num_combinations <= 0.0) return "player 0 has an empty range";
  spec.root_street = (Street)street;
  spec.board.assign(c.board, c.board + n);
  Betting& b = spec.root;
  b.street = (Street)c.street;
  b.pot = c.pot;
  b.commit[0] = c.commit[0]; b.commit[1] = c.commit[1];
  b.stack[0] = c.stack[0]; b.stack[1] = c.stack[1];
  b.current_bet = c.current_bet;
  b.last_raise_inc = c.last_raise_inc;
  b.num_raises = c.num_raises;
  b.to_act = c.to_act;
  b.needs[0] = c.needs[0] != 0; b.needs[1] = c.needs[1] != 0;
  spec.menu = menu_of(c.menu);
  spec.future_menu = c.has_future_menu ? menu_of(c.future_menu) : spec.menu;
  spec.big_blind = c.big_blind;
  spec.rake_rate = c.rake_rate;
  spec.rake_cap = c.rake_cap;
  spec.hands[0] = hands.n(0);
  spec.hands[1] = hands.n(1);
  return "";
}

// tree.rs:264-272: storage bytes per (action, hand) value pair (8 f32, 4
// 16-bit) plus the arena's 56-byte nodes
static uint64_t memory_bytes(const TreeCounts& t, bool compressed = false) {
  return t.cells * (compressed ? 4 : 8) + (t.player + t.chance + t.terminal) * 56;
}
// lib.rs:346-352: Compressed, or Auto when the f32 tree would exceed 1 GB
static bool compressed_for(const pts_config_c& c, const TreeCounts& t) {
  return c.storage_mode == 1 || (c.storage_mode == 2 && memory_bytes(t) > (1ull << 30));
}

// ------------------------------------------------------------------ engine
extern "C" {

size_t pts_abi_sizeof(int what) {
  switch (what) {
    case 0: return sizeof(pts_config_c);
    case 1: return sizeof(pts_tree_desc);
    case 2: return sizeof(pts_progress_c);
    case 3: return sizeof(pts_step_c);
    case 4: return sizeof(pts_stats_c);
    default: return 0;
  }
}

void pts_last_error_detail(int32_t* hip_error, uint64_t* needed, uint64_t* available) {
  const ErrorDetail& d = last_error_detail();
  if (hip_error) *hip_error = d.hip_error;
  if (needed) *needed = d.needed;
  if (available) *available = d.available;
}

size_t pts_last_error(char* buf, size_t len) {
  const std::string& e = last_error();
  if (buf && len) std::snprintf(buf, len, "%s", e.c_str());
  return e.size();
}

pts_status pts_evaluate(const uint8_t* cards, uint32_t ncards, uint64_t count, uint32_t* out) {  // eval.rs:142-198
  if (ncards < 5 || ncards > 7) return bad("invalid input: a hand has 5 to 7 cards");
  if (count && (!cards || !out)) return bad("invalid input: null buffer");
  for (uint64_t i = 0; i < count; i++) {
    const uint8_t* c = cards + i * ncards;
    uint64_t seen = 0, lanes = 0;
    for (uint32_t k = 0; k < ncards; k++) {
      if (c[k] >= 52 || (seen >> c[k] & 1)) return bad("invalid input: hand " + std::to_string(i) + " repeats a card or names one above 51");
      seen |= 1ull << c[k];
      lanes |= lane_bit(c[k]);
    }
    out[i] = hand_strength(lanes);
  }
  return PTS_OK;
}

pts_status pts_device_count(int* out) {
  *out = 0;
  hipError_t e = hipGetDeviceCount(out);
  if (e != hipSuccess) {
    *out = 0;
    set_error(std::string("no GPU: ") + hipGetErrorString(e));
    set_error_detail(ErrorDetail{(int32_t)e, 0, 0});
    return PTS_NO_DEVICE;
  }
  return PTS_OK;
}

// lib.rs:496-503 keeps river roots on the CPU under Auto (an RTX 4060's GPU
// solve of a river tree lost to the CPU). On MI355X a river root solves to
// 0.5% of the pot faster on the GPU than on the host CPU, tree build included
// (bench.py --workload river-auto), so Auto puts every street on the GPU.
int pts_backend_wants_gpu(int backend, int street) {
  if (backend == PTS_BACKEND_GPU) return 1;
  if (backend == PTS_BACKEND_AUTO) return street >= kFlop && street <= kRiver;
  return 0;
}

static PlanInput desc_input(const pts_tree_desc* t) {
  PlanInput in;
  in.num_nodes = t->num_nodes;
  in.kind = t->kind;
  in.street = t->street;
  in.num_children = t->num_children;
  in.first_child = t->first_child;
  in.invest = t->invest;
  in.offset = t->offset;
  in.rake = t->rake;
  in.chance_factor = t->chance_factor;
  in.card = t->card;
  in.num_boards = t->num_boards;
  in.boards = t->boards;
  for (int q = 0; q < 2; q++) {
    in.num_hands[q] = t->num_hands[q];
    in.hand_cards[q] = t->hand_cards[q];
    in.hand_weights[q] = t->hand_weights[q];
    in.hand_same[q] = t->hand_same[q];
  }
  in.half_pot = t->half_pot;
  in.num_combinations = t->num_combinations;
  in.storage_len = t->storage_len;
  in.batch_river = t->batch_river != 0;
  return in;
}

pts_status pts_engine_create(int device, const pts_tree_desc* t, const float* regrets, const float* sums, int precision,
                             uint64_t scratch_budget, pts_engine** out) {
  *out = nullptr;
  if (!t) return bad("invalid input: null argument");  // null regrets or sums: zeros
  PlanInput in = desc_input(t);
  Plan plan;
  std::string err = build_plan(in, plan);
  if (!err.empty()) return bad(err);
  std::unique_ptr<Engine> e;
  Status s = Engine::create(device, std::move(plan), regrets, sums, precision, scratch_budget, e);
  if (s != kOk) return st(s);
  *out = new pts_engine{std::move(e)};
  return PTS_OK;
}

pts_status pts_engine_create_q(int device, const pts_tree_desc* t, const uint32_t* slot, uint32_t num_slots,
                               const int16_t* regrets, const uint16_t* sums, const float* regret_scale,
                               const float* strat_scale, int precision, uint64_t scratch_budget, pts_engine** out) {
  *out = nullptr;
  if (!t || !slot) return bad("invalid input: null argument");
  const bool zeros = !regrets && !sums && !regret_scale && !strat_scale;
  if (!zeros && (!regrets || !sums || !regret_scale || !strat_scale))
    return bad("invalid input: give all four storage arrays, or none for zeros");
  if (num_slots == 0) return bad("invalid input: no player slots");
  PlanInput in = desc_input(t);
  in.pslot = slot;
  in.num_pslots = num_slots;
  Plan plan;
  std::string err = build_plan(in, plan);
  if (!err.empty()) return bad(err);
  std::unique_ptr<Engine> e;
  Status s = Engine::create(device, std::move(plan), nullptr, nullptr, precision, scratch_budget, e, true);
  if (s != kOk) return st(s);
  if (!zeros && (s = e->upload_q(regrets, sums, regret_scale, strat_scale)) != kOk) return st(s);
  *out = new pts_engine{std::move(e)};
  return PTS_OK;
}

pts_status pts_engine_download_q(pts_engine* e, int16_t* regrets, uint16_t* sums, float* regret_scale, float* strat_scale) {
  return st(e->e->download_q(regrets, sums, regret_scale, strat_scale));
}

pts_status pts_engine_iterate(pts_engine* e, float pos, float neg, float strat, int rw) {
  return st(e->e->cfr_iteration({pos, neg, strat, rw}));
}

pts_status pts_engine_traverse(pts_engine* e, int player, int mode, float pos, float neg, float strat, int rw, double* out,
                               size_t n) {
  if (player < 0 || player > 1 || mode < 0 || mode > 2) return bad("invalid input: player or mode");
  if (n != e->e->plan().n[player]) return bad("invalid input: output length");
  std::vector<double> v;
  Status s = e->e->traverse(player, (Mode)mode, {pos, neg, strat, rw}, v);
  if (s != kOk) return st(s);
  std::memcpy(out, v.data(), n * sizeof(double));
  return PTS_OK;
}

pts_status pts_engine_root_values(pts_engine* e, int mode, double out[2]) {
  if (mode != PTS_MODE_BEST_RESPONSE && mode != PTS_MODE_AVERAGE) return bad("invalid input: mode");
  return st(e->e->root_values((Mode)mode, out));
}

pts_status pts_engine_exploitability(pts_engine* e, int raked, double* out) { return st(e->e->exploitability(raked != 0, out)); }

pts_status pts_engine_download(pts_engine* e, float* regrets, float* sums, uint64_t len) {
  if (len != e->e->plan().storage_len) return bad("invalid input: storage length");
  return st(e->e->download(regrets, sums));
}

pts_status pts_engine_upload(pts_engine* e, const float* regrets, const float* sums, uint64_t len) {
  if (len != e->e->plan().storage_len) return bad("invalid input: storage length");
  return st(e->e->upload(regrets, sums));
}

pts_status pts_engine_subtree_values(pts_engine* e, uint32_t node, int player, const float* opp, size_t nq, float* out, size_t len) {
  const Plan& P = e->e->plan();
  if (player < 0 || player > 1 || node >= P.num_nodes) return bad("invalid input: node or player");
  if (len != (size_t)P.na[node] * P.n[player]) return bad("invalid input: output length");
  std::vector<float> r(opp, opp + nq), v;
  Status s = e->e->subtree_values(node, player, r, v);
  if (s != kOk) return st(s);
  std::memcpy(out, v.data(), len * sizeof(float));
  return PTS_OK;
}

uint64_t pts_engine_device_bytes(pts_engine* e) { return e->e->device_bytes(); }
uint32_t pts_engine_num_groups(pts_engine* e, int f64) { return (uint32_t)e->e->groups(f64 != 0).size(); }
uint64_t pts_engine_jit_segments(pts_engine* e) { return e->e->jit_segments(); }

pts_status pts_engine_profile_iteration(pts_engine* e, float pos, float neg, float strat, int rw, double ms[4],
                                        uint32_t launches[4], double* wall_ms) {
  Profile p;
  Status s = e->e->profile_iteration({pos, neg, strat, rw}, p);
  if (s != kOk) return st(s);
  for (int k = 0; k < 4; k++) {
    ms[k] = p.ms[k];
    launches[k] = p.launches[k];
  }
  if (wall_ms) *wall_ms = p.wall_ms;
  return PTS_OK;
}
void pts_engine_destroy(pts_engine* e) { delete e; }

// ------------------------------------------------------------------ solver
pts_status pts_solver_estimate(const pts_config_c* c, pts_stats_c* out) {  // lib.rs:355-362
  TreeSpec spec;
  Hands hands;
  std::string err = validate(*c, spec, hands);
  if (!err.empty()) return bad(err);
  const Iso iso = make_iso(hands, c->isomorphism != 0);
  spec.iso = &iso;
  TreeCounts t = count_tree(spec);
  *out = {t.player, t.chance, t.terminal, t.cells, memory_bytes(t, compressed_for(*c, t))};
  return PTS_OK;
}

static PlanInput plan_input(pts_solver* s, std::vector<uint8_t>& kind, std::vector<uint8_t>& street, std::vector<uint16_t>& nc,
                            std::vector<uint32_t>& fc, std::vector<int32_t>& inv, std::vector<uint64_t>& off,
                            std::vector<float>& rake, std::vector<float>& fac, std::vector<int8_t>& card,
                            std::vector<int32_t>& boards, std::vector<uint32_t>* same, std::vector<uint32_t>& isf,
                            std::vector<uint16_t>& isc, std::vector<uint16_t>* perm, std::vector<uint32_t>& ps) {
  size_t N = s->tree.nodes.size();
  kind.resize(N); street.resize(N); nc.resize(N); fc.resize(N); inv.resize(2 * N); off.resize(N);
  rake.resize(N); fac.resize(N); card.resize(N); isf.resize(N); isc.resize(N); ps.assign(N, 0);
  for (size_t i = 0; i < N; i++) {
    const Node& n = s->tree.nodes[i];
    kind[i] = n.kind; street[i] = n.street; nc[i] = n.num_children; fc[i] = n.first_child;
    inv[2 * i] = n.invest[0]; inv[2 * i + 1] = n.invest[1]; off[i] = n.offset;
    rake[i] = n.rake; fac[i] = n.chance_factor; card[i] = n.card;
    isf[i] = n.iso_first; isc[i] = n.iso_cnt;
    if (is_player(n.kind)) ps[i] = n.pslot;
  }
  for (auto& b : s->tree.boards) boards.insert(boards.end(), b.begin(), b.end());
  PlanInput in;
  in.num_nodes = N;
  in.kind = kind.data(); in.street = street.data(); in.num_children = nc.data(); in.first_child = fc.data();
  in.invest = inv.data(); in.offset = off.data(); in.rake = rake.data(); in.chance_factor = fac.data(); in.card = card.data();
  in.num_boards = (uint32_t)s->tree.boards.size();
  in.boards = boards.data();
  for (int q = 0; q < 2; q++) {
    in.num_hands[q] = (uint32_t)s->hands.n(q);
    in.hand_cards[q] = s->hands.cards[q].data();
    in.hand_weights[q] = s->hands.weights[q].data();
    same[q] = s->hands.same[q];
    in.hand_same[q] = same[q].data();
  }
  in.half_pot = (double)s->cfg.pot / 2.0;
  in.num_combinations = s->hands.num_combinations;
  in.storage_len = s->tree.storage_len;
  in.pslot = ps.data();
  in.num_pslots = s->tree.player_nodes;
  in.batch_river = true;
  in.root_stack[0] = s->cfg.stack[0];
  in.root_stack[1] = s->cfg.stack[1];
  if (!s->tree.iso_refs.empty()) {
    // every pair's permutation, identity where the pair is not symmetric
    for (int q = 0; q < 2; q++) {
      const size_t n = s->hands.n(q);
      perm[q].resize(6 * n);
      for (int pi = 0; pi < 6; pi++)
        for (size_t h = 0; h < n; h++) perm[q][pi * n + h] = s->iso.symmetric[pi] ? s->iso.perm[q][pi][h] : (uint16_t)h;
      in.iso_perm[q] = perm[q].data();
    }
    in.iso_first = isf.data();
    in.iso_cnt = isc.data();
    in.iso_refs = s->tree.iso_refs.data();
  }
  return in;
}

}  // extern "C"

// Solver::new on the host (lib.rs:365-396): validation, the tree, the hands
// and the validated GPU plan.
static std::string host_build(const pts_config_c* c, pts_solver* s, Plan& plan, bool depth_limit = false) {
  s->cfg = *c;
  std::string err = validate(*c, s->spec, s->hands);
  if (!err.empty()) return err;
  s->spec.depth_limit = depth_limit;
  s->iso = make_iso(s->hands, c->isomorphism != 0);  // iso.rs:245-273
  s->spec.iso = &s->iso;
  s->counts = count_tree(s->spec);
  s->compressed = compressed_for(*c, s->counts);
  uint64_t need = memory_bytes(s->counts, s->compressed);
  if (need > c->max_memory_bytes) {
    char m[160];
    std::snprintf(m, sizeof m, "game tree needs ~%.2f GB but the limit is %.2f GB", (double)need / 1e9,
                  (double)c->max_memory_bytes / 1e9);
    return m;
  }
  s->tree = build_tree(s->spec);
  std::vector<uint8_t> kind, street;
  std::vector<uint16_t> nc;
  std::vector<uint32_t> fc;
  std::vector<int32_t> inv, boards;
  std::vector<uint64_t> off;
  std::vector<float> rake, fac;
  std::vector<int8_t> card;
  std::vector<uint32_t> same[2], isf, ps;
  std::vector<uint16_t> isc, perm[2];
  PlanInput in = plan_input(s, kind, street, nc, fc, inv, off, rake, fac, card, boards, same, isf, isc, perm, ps);
  return build_plan(in, plan);
}

extern "C" {

pts_status pts_solver_new(const pts_config_c* c, int device, int precision, uint64_t scratch_budget,
                          pts_solver** out) {  // lib.rs:365-396
  *out = nullptr;
  std::unique_ptr<pts_solver> s(new pts_solver());
  Plan plan;
  std::string err = host_build(c, s.get(), plan);
  if (!err.empty()) return bad(err);
  if (device < 0) {  // host only: tree, hands and a validated plan, no engine
    s->host_plan = std::move(plan);
    *out = s.release();
    return PTS_OK;
  }
  Status e = Engine::create(device, std::move(plan), nullptr, nullptr, precision, scratch_budget, s->engine, s->compressed);
  if (e != kOk) return st(e);
  *out = s.release();
  return PTS_OK;
}

void pts_solver_free(pts_solver* s) { delete s; }

void pts_solver_stats(pts_solver* s, pts_stats_c* out) {
  const TreeCounts& t = s->counts;
  *out = {t.player, t.chance, t.terminal, t.cells, memory_bytes(t, s->compressed)};
}

int32_t pts_solver_is_compressed(pts_solver* s) { return s->compressed ? 1 : 0; }

static pts_status need_engine(pts_solver* s) {
  if (!s->engine) {
    set_error("no GPU engine (the solver was built host-only)");
    return PTS_NO_DEVICE;
  }
  return PTS_OK;
}
#define NEED_ENGINE(s) do { pts_status r_ = need_engine(s); if (r_ != PTS_OK) return r_; } while (0)

static bool raked(const pts_solver* s) { return s->cfg.rake_rate > 0.0 && s->cfg.rake_cap > 0.0; }

pts_status pts_run_solve(uint32_t* iterations, double* exploitability, double pot, uint32_t max_iterations,
                         double target_pct, uint32_t check_every, const uint8_t* cancel, pts_iterate_fn iterate,
                         pts_measure_fn measure, pts_progress_fn on_progress, void* user, pts_progress_c* out) {
  // crates/solver/src/lib.rs:594-642 (run_solve)
  if (!iterations || !exploitability || !iterate || !measure || !out) return bad("invalid input: null argument");
  if (!(pot > 0.0)) return bad("invalid input: the pot must be positive");
  const uint32_t ce = std::max(check_every, 1u);
  pts_progress_c pr{*iterations, max_iterations, *exploitability, 100.0 * *exploitability / pot, 0, 0};
  const double target = target_pct * pot / 100.0;
  uint32_t next_check = *iterations + ce;
  bool has_last = false;
  uint32_t lt = 0;
  double le = 0.0;
  while (*iterations < max_iterations) {
    if (cancel && __atomic_load_n(cancel, __ATOMIC_RELAXED)) {
      pr.finished = 1;
      *out = pr;
      set_error("cancelled");
      return PTS_CANCELLED;
    }
    *iterations += 1;
    int32_t rc = iterate(user, *iterations);
    if (rc != PTS_OK) return (pts_status)rc;
    const bool last = *iterations >= max_iterations;
    if (*iterations >= next_check || last) {
      double x = 0.0;
      if ((rc = measure(user, &x)) != PTS_OK) return (pts_status)rc;
      *exploitability = x;
      pr.iteration = *iterations;
      pr.exploitability = x;
      pr.exploitability_pct = 100.0 * x / pot;
      const bool done = pr.exploitability_pct <= target_pct || last;
      pr.finished = done;
      const bool keep_going = on_progress ? on_progress(&pr, user) != 0 : true;
      if (done || !keep_going) {
        pr.finished = 1;
        *out = pr;
        if (!done) {
          set_error("cancelled by the progress callback");
          return PTS_CANCELLED;
        }
        return PTS_OK;
      }
      const uint32_t t = *iterations;
      next_check = t + check_gap(has_last, lt, le, t, x, target, ce);
      has_last = true;
      lt = t;
      le = x;
    }
  }
  pr.finished = 1;
  *out = pr;
  return PTS_OK;
}

namespace {
// The solver's engine under pts_run_solve, with the caller's callback and a
// history of the checks.
struct SolveCtx {
  pts_solver* s;
  pts_progress_fn user_fn;
  void* user;
  pts_progress_c* history;
  int32_t cap, len;
};
int32_t solve_iterate(void* u, uint32_t t) {
  SolveCtx* c = (SolveCtx*)u;
  return (int32_t)st(c->s->engine->cfr_iteration(discount(t, c->s->cfg)));
}
int32_t solve_measure(void* u, double* x) {
  SolveCtx* c = (SolveCtx*)u;
  const bool raked = c->s->cfg.rake_rate > 0.0 && c->s->cfg.rake_cap > 0.0;
  return (int32_t)st(c->s->engine->exploitability(raked, x));
}
int32_t solve_progress(const pts_progress_c* p, void* u) {
  SolveCtx* c = (SolveCtx*)u;
  if (c->history && c->len < c->cap) c->history[c->len] = *p;
  c->len++;
  return c->user_fn ? c->user_fn(p, c->user) : 1;
}
}  // namespace

static pts_status solver_solve(pts_solver* s, uint32_t max_iterations, double target_pct, uint32_t check_every,
                               const uint8_t* cancel, pts_progress_fn on_progress, void* user, pts_progress_c* out,
                               pts_progress_c* history, int32_t history_cap, int32_t* history_len) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  SolveCtx c{s, on_progress, user, history, std::max(history_cap, 0), 0};
  pts_status r = pts_run_solve(&s->iterations, &s->exploitability, (double)s->cfg.pot, max_iterations, target_pct,
                               check_every, cancel, solve_iterate, solve_measure, solve_progress, &c, out);
  if (history_len) *history_len = c.len;
  return r;
}

pts_status pts_solver_solve(pts_solver* s, uint32_t max_iterations, double target_pct, uint32_t check_every, pts_progress_c* out,
                            pts_progress_c* history, int32_t history_cap, int32_t* history_len) {  // lib.rs:475-478
  return solver_solve(s, max_iterations, target_pct, check_every, nullptr, nullptr, nullptr, out, history, history_cap,
                      history_len);
}

pts_status pts_solver_solve_cancellable(pts_solver* s, uint32_t max_iterations, double target_pct, uint32_t check_every,
                                        const uint8_t* cancel, pts_progress_fn on_progress, void* user,
                                        pts_progress_c* out) {  // lib.rs:480-494
  return solver_solve(s, max_iterations, target_pct, check_every, cancel, on_progress, user, out, nullptr, 0, nullptr);
}

pts_status pts_solver_iterate(pts_solver* s, uint32_t count) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  for (uint32_t i = 0; i < count; i++) {
    s->iterations += 1;
    Status e = s->engine->cfr_iteration(discount(s->iterations, s->cfg));
    if (e != kOk) return st(e);
  }
  return PTS_OK;
}

pts_status pts_solver_exploitability(pts_solver* s, double* out) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  return st(s->engine->exploitability(raked(s), out));
}
pts_status pts_solver_root_values(pts_solver* s, int mode, double out[2]) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  if (mode != PTS_MODE_BEST_RESPONSE && mode != PTS_MODE_AVERAGE) return bad("invalid input: mode");
  return st(s->engine->root_values((Mode)mode, out));
}
uint32_t pts_solver_iterations(pts_solver* s) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  return s->iterations;
}
void pts_solver_set_iterations(pts_solver* s, uint32_t t) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  s->iterations = t;
}
uint64_t pts_solver_storage_len(pts_solver* s) { return s->tree.storage_len; }
uint32_t pts_solver_num_groups(pts_solver* s, int f64) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  return s->engine ? (uint32_t)s->engine->groups(f64 != 0).size() : 0u;
}
uint64_t pts_solver_jit_segments(pts_solver* s) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  return s->engine ? s->engine->jit_segments() : 0u;
}

pts_status pts_solver_plan_groups(pts_solver* s, uint64_t scratch_budget, int f64, uint32_t* num_groups,
                                  uint64_t* scratch_bytes) {
  if (s->engine) return bad("invalid input: the solver has an engine (use pts_solver_num_groups)");
  const Plan& P = s->host_plan;
  const uint32_t rowlen = (std::max(P.n[0], P.n[1]) + 31u) & ~31u;
  std::vector<ChunkGroup> g;
  ScratchRows r = partition_chunks(P, f64 ? 8 : 4, rowlen, scratch_budget, g);
  std::string err = check_groups(P, g, r);
  if (!err.empty()) return bad("invalid input: " + err);
  *num_groups = (uint32_t)g.size();
  *scratch_bytes = r.reach_bytes + r.val_bytes + r.alive_rows * 4;
  return PTS_OK;
}
int32_t pts_solver_num_hands(pts_solver* s, int p) { return (int32_t)s->hands.n(p); }
// f32 storage as is; 16-bit storage decoded (storage.rs:209-246)
pts_status pts_solver_get_storage(pts_solver* s, float* r, float* m) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  if (!s->compressed) return st(s->engine->download(r, m));
  const size_t len = s->tree.storage_len, k = s->tree.player_nodes;
  std::vector<int16_t> rq(len);
  std::vector<uint16_t> sq(len);
  std::vector<float> rs(k), ss(k);
  Status e = s->engine->download_q(rq.data(), sq.data(), rs.data(), ss.data());
  if (e != kOk) return st(e);
  for (const Node& n : s->tree.nodes) {
    if (!is_player(n.kind)) continue;
    const size_t cells = (size_t)n.num_children * s->hands.n(n.kind);
    const float rstep = rs[n.pslot] / 32767.0f, sstep = ss[n.pslot] / 65535.0f;
    for (size_t i = n.offset; i < n.offset + cells; i++) {
      if (r) r[i] = (float)rq[i] * rstep;
      if (m) m[i] = (float)sq[i] * sstep;
    }
  }
  return PTS_OK;
}
pts_status pts_solver_set_storage(pts_solver* s, const float* r, const float* m) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  if (s->compressed) return bad("invalid input: 16-bit storage is set through pts_solver_set_storage_q");
  return st(s->engine->upload(r, m));
}
pts_status pts_solver_get_storage_q(pts_solver* s, int16_t* r, uint16_t* m, float* regret_scale, float* strat_scale) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  if (!s->compressed) return bad("invalid input: the solver has f32 storage");
  return st(s->engine->download_q(r, m, regret_scale, strat_scale));
}
pts_status pts_solver_set_storage_q(pts_solver* s, const int16_t* r, const uint16_t* m, const float* regret_scale,
                                    const float* strat_scale) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  if (!s->compressed) return bad("invalid input: the solver has f32 storage");
  return st(s->engine->upload_q(r, m, regret_scale, strat_scale));
}
pts_status pts_solver_profile_iteration(pts_solver* s, double ms[4], uint32_t launches[4], double* wall_ms) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  NEED_ENGINE(s);
  s->iterations += 1;
  IterParams d = discount(s->iterations, s->cfg);
  Profile p;
  Status e = s->engine->profile_iteration(d, p);
  if (e != kOk) return st(e);
  for (int k = 0; k < 4; k++) {
    ms[k] = p.ms[k];
    launches[k] = p.launches[k];
  }
  if (wall_ms) *wall_ms = p.wall_ms;
  return PTS_OK;
}
uint64_t pts_solver_num_nodes(pts_solver* s) { return s->tree.nodes.size(); }

void pts_solver_nodes(pts_solver* s, uint8_t* kind, uint8_t* street, uint16_t* nc, uint32_t* fc, int32_t* inv, uint64_t* off,
                      float* rake, float* fac, int8_t* card, int32_t* ak, int64_t* to) {
  for (size_t i = 0; i < s->tree.nodes.size(); i++) {
    const Node& n = s->tree.nodes[i];
    kind[i] = n.kind; street[i] = n.street; nc[i] = n.num_children; fc[i] = n.first_child;
    inv[2 * i] = n.invest[0]; inv[2 * i + 1] = n.invest[1]; off[i] = n.offset;
    rake[i] = n.rake; fac[i] = n.chance_factor; card[i] = n.card;
    bool act_edge = i > 0 && n.card < 0;
    ak[i] = act_edge ? (int32_t)n.action.kind : -1;
    to[i] = act_edge && n.action.kind == kRaise ? n.action.to : 0;
  }
}

// lib.rs:746-825: reach of both players along `path`. A card merged by suit
// isomorphism continues at its canonical child: later cards are read through
// the swapped suits, the reach arrays are permuted into the canonical hands,
// and to_canon maps each hand to the canonical hand that stands for it.
static pts_status walk(pts_solver* s, const pts_step_c* path, int32_t np, uint32_t& id, std::vector<float> reach[2],
                       std::vector<uint32_t> to_canon[2]) {
  NEED_ENGINE(s);
  id = 0;
  reach[0] = s->hands.weights[0];
  reach[1] = s->hands.weights[1];
  int suit_map[4] = {0, 1, 2, 3};
  for (int x = 0; x < 2; x++) {
    to_canon[x].resize(s->hands.n(x));
    for (size_t h = 0; h < to_canon[x].size(); h++) to_canon[x][h] = (uint32_t)h;
  }
  for (int32_t d = 0; d < np; d++) {
    const Node& node = s->tree.nodes[id];
    const pts_step_c& stp = path[d];
    if (is_player(node.kind) && stp.kind == 0) {
      int ai = -1;
      for (int a = 0; a < node.num_children; a++) {
        const Action& x = s->tree.nodes[node.first_child + a].action;
        if ((int)x.kind == stp.act_kind && (x.kind != kRaise || x.to == stp.to)) { ai = a; break; }
      }
      if (ai < 0) return bad("bad path: action not available at depth " + std::to_string(d));
      int q = node.kind;
      size_t n = s->hands.n(q);
      std::vector<float> strat((size_t)node.num_children * n);
      Status e = s->engine->read_sums(node.offset, strat.size(), strat.data(), node.pslot);
      if (e != kOk) return st(e);
      normalize(strat.data(), node.num_children, n);
      for (size_t i = 0; i < n; i++) reach[q][i] *= strat[(size_t)ai * n + i];
      id = node.first_child + ai;
    } else if (node.kind == kChance && stp.kind == 1) {
      if (stp.card < 0 || stp.card >= 52) return bad("bad path: bad card at depth " + std::to_string(d));
      const int cc = (stp.card >> 2) * 4 + suit_map[stp.card & 3];
      uint32_t ch = kNone;
      for (uint32_t c = node.first_child; c < node.first_child + node.num_children; c++)
        if (s->tree.nodes[c].card == cc) { ch = c; break; }
      if (ch == kNone) {
        const IsoRef* e = nullptr;
        for (uint32_t k = 0; k < node.iso_cnt; k++)
          if (s->tree.iso_refs[node.iso_first + k].card == cc) e = &s->tree.iso_refs[node.iso_first + k];
        if (!e) return bad("bad path: card not in the tree at depth " + std::to_string(d));
        const int a = kSuitPairs[e->pair][0], b = kSuitPairs[e->pair][1];
        for (int& x : suit_map) x = x == a ? b : x == b ? a : x;
        for (int x = 0; x < 2; x++) {
          const std::vector<uint16_t>& perm = s->iso.perm[x][e->pair];
          std::vector<float> r2(reach[x].size(), 0.0f);
          for (size_t h = 0; h < reach[x].size(); h++) r2[perm[h]] = reach[x][h];
          reach[x] = std::move(r2);
          for (uint32_t& t : to_canon[x]) t = perm[t];
        }
        ch = node.first_child + e->child;
      }
      id = ch;
      const int canonical = s->tree.nodes[ch].card;
      for (int x = 0; x < 2; x++)
        for (size_t h = 0; h < s->hands.n(x); h++)
          if (s->hands.cards[x][2 * h] == canonical || s->hands.cards[x][2 * h + 1] == canonical) reach[x][h] = 0.0f;
    } else {
      return bad("bad path: step doesn't match node kind at depth " + std::to_string(d));
    }
  }
  return PTS_OK;
}

pts_status pts_solver_query(pts_solver* s, const pts_step_c* path, int32_t npath, int32_t* player, int32_t* na, int32_t* n,
                            int32_t* act_kinds, int64_t* act_tos, uint16_t* combos, float* strategy, float* ev, float* reach_out,
                            double reach_fraction[2], int64_t* pot, int64_t invest[2], int32_t cap_actions, int32_t cap_hands) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  uint32_t id;
  std::vector<float> reach[2];
  std::vector<uint32_t> to_canon[2];
  pts_status r = walk(s, path, npath, id, reach, to_canon);
  if (r != PTS_OK) return r;
  const Node node = s->tree.nodes[id];
  if (!is_player(node.kind)) return bad("bad path: not a decision node");
  int q = node.kind, o = q ^ 1;
  size_t nh = s->hands.n(q), A = node.num_children;
  if ((int32_t)A > cap_actions || (int32_t)nh > cap_hands) return bad("invalid input: capacity too small");
  std::vector<double> compat = compatible_reach(s->hands, q, reach[o].data());
  std::vector<float> cv;
  Status e = s->engine->subtree_values(id, q, reach[o], cv);
  if (e != kOk) return st(e);
  std::vector<float> strat(A * nh);
  if ((e = s->engine->read_sums(node.offset, strat.size(), strat.data(), node.pslot)) != kOk) return st(e);
  normalize(strat.data(), A, nh);
  double half_pot = (double)s->cfg.pot / 2.0, norm = 1.0 / s->hands.num_combinations;
  double offset = half_pot + (double)node.invest[q];
  for (size_t h = 0; h < nh; h++) {
    const size_t hc = to_canon[q][h];  // the canonical hand standing for h
    reach_out[h] = reach[q][hc];
    double denom = compat[hc] * norm;
    for (size_t a = 0; a < A; a++) {
      strategy[a * nh + h] = strat[a * nh + hc];
      ev[a * nh + h] = denom > 0.0 ? (float)((double)cv[a * nh + hc] / denom + offset) : 0.0f;
    }
  }
  for (int p = 0; p < 2; p++) {
    double root = 0.0, now = 0.0;
    for (float x : s->hands.weights[p]) root += (double)x;
    for (float x : reach[p]) now += (double)x;
    reach_fraction[p] = root > 0.0 ? now / root : 0.0;
  }
  *player = q;
  *na = (int32_t)A;
  *n = (int32_t)nh;
  for (size_t a = 0; a < A; a++) {
    const Action& x = s->tree.nodes[node.first_child + a].action;
    act_kinds[a] = x.kind;
    act_tos[a] = x.kind == kRaise ? x.to : 0;
  }
  std::memcpy(combos, s->hands.combos[q].data(), nh * sizeof(uint16_t));
  *pot = s->cfg.pot + node.invest[0] + node.invest[1];
  invest[0] = node.invest[0];
  invest[1] = node.invest[1];
  return PTS_OK;
}

pts_status pts_solver_ranges_after(pts_solver* s, const pts_step_c* path, int32_t npath, float* out0, float* out1) {
  std::lock_guard<std::recursive_mutex> lock_(s->mu);
  uint32_t id;
  std::vector<float> reach[2];
  std::vector<uint32_t> to_canon[2];
  pts_status r = walk(s, path, npath, id, reach, to_canon);
  if (r != PTS_OK) return r;
  float* outs[2] = {out0, out1};
  for (int p = 0; p < 2; p++) {
    std::memset(outs[p], 0, 1326 * sizeof(float));
    for (size_t h = 0; h < s->hands.n(p); h++) outs[p][s->hands.combos[p][h]] = reach[p][to_canon[p][h]];
  }
  return PTS_OK;
}

// ------------------------------------------------------------------ batch
}  // extern "C"

struct pts_batch {
  std::vector<std::unique_ptr<pts_solver>> spots;  // host side of every spot (no engines)
  std::vector<uint64_t> store_off;
  std::unique_ptr<Engine> engine;
  uint32_t iterations = 0;
  std::vector<uint32_t> active;
  Plan host_plan;                  // host-only batches (device -1)
  uint64_t host_scratch[2] = {0, 0};
};

#define NEED_BATCH_ENGINE(b) \
  if (!(b)->engine) return bad("no GPU engine (the batch was built host-only)")

static bool same_dcfr(const pts_config_c& a, const pts_config_c& b) {
  return a.alpha == b.alpha && a.beta_some == b.beta_some && a.beta == b.beta && a.gamma == b.gamma &&
         a.restart_average == b.restart_average && a.unweighted_average == b.unweighted_average;
}

// Exploitability (chips) of the spots `mask` names (the others' traversals
// are skipped and their entries left 0), then the solve's own active set back.
static pts_status batch_expl(pts_batch* b, const std::vector<uint32_t>& mask, std::vector<double>& out) {
  const uint32_t K = (uint32_t)b->spots.size();
  Status e;
  if ((e = b->engine->set_active(mask)) != kOk) return st(e);
  std::vector<double> br, ev;
  if ((e = b->engine->root_values_all(kBest, br)) != kOk) return st(e);
  bool any_raked = false;
  for (auto& s : b->spots) any_raked |= raked(s.get());
  if (any_raked && (e = b->engine->root_values_all(kAverage, ev)) != kOk) return st(e);
  out.assign(K, 0.0);
  for (uint32_t k = 0; k < K; k++)  // cfr.rs:494-502
    if (mask[k])
      out[k] = raked(b->spots[k].get()) ? ((br[2 * k] - ev[2 * k]) + (br[2 * k + 1] - ev[2 * k + 1])) / 2.0
                                      : (br[2 * k] + br[2 * k + 1]) / 2.0;
  if ((e = b->engine->set_active(b->active)) != kOk) return st(e);
  return PTS_OK;
}

extern "C" {

pts_status pts_batch_new(const pts_config_c* cfgs, uint32_t num_spots, int device, int precision, pts_batch** out) {
  *out = nullptr;
  if (num_spots == 0) return bad("invalid input: empty batch");
  std::unique_ptr<pts_batch> b(new pts_batch());
  std::vector<Plan> plans(num_spots);
  uint64_t off = 0;
  for (uint32_t k = 0; k < num_spots; k++) {
    if (!same_dcfr(cfgs[k], cfgs[0])) return bad("invalid input: the spots of a batch share one set of DCFR parameters");
    std::unique_ptr<pts_solver> s(new pts_solver());
    std::string err = host_build(&cfgs[k], s.get(), plans[k]);
    if (!err.empty()) return bad("spot " + std::to_string(k) + ": " + err);
    if (s->compressed) return bad("spot " + std::to_string(k) + ": batched spots need f32 storage");
    b->store_off.push_back(off);
    off += s->tree.storage_len;
    b->spots.push_back(std::move(s));
  }
  Plan merged;
  std::string err = merge_plans(std::move(plans), merged);
  if (!err.empty()) return bad(err);
  b->active.assign(num_spots, 1u);
  if (device < 0) {  // host only: the merged plan, checked as an engine would check it
    const uint32_t rowlen = (std::max(merged.n[0], merged.n[1]) + 31u) & ~31u;
    for (int f64 = 0; f64 < 2; f64++) {
      std::vector<ChunkGroup> g;
      ScratchRows r = partition_chunks(merged, f64 ? 8 : 4, rowlen, ~0ull, g);
      err = check_groups(merged, g, r);
      if (!err.empty()) return bad("invalid input: " + err);
      b->host_scratch[f64] = r.reach_bytes + r.val_bytes + r.alive_rows * 4;
    }
    b->host_plan = std::move(merged);
    *out = b.release();
    return PTS_OK;
  }
  Status e = Engine::create(device, std::move(merged), nullptr, nullptr, precision, 0, b->engine);
  if (e != kOk) return st(e);
  *out = b.release();
  return PTS_OK;
}

void pts_batch_free(pts_batch* b) { delete b; }
uint32_t pts_batch_num_spots(pts_batch* b) { return (uint32_t)b->spots.size(); }
uint32_t pts_batch_iterations(pts_batch* b) { return b->iterations; }

pts_status pts_batch_stats(pts_batch* b, uint32_t spot, pts_stats_c* out) {
  if (spot >= b->spots.size()) return bad("invalid input: spot");
  const TreeCounts& t = b->spots[spot]->counts;
  *out = {t.player, t.chance, t.terminal, t.cells, memory_bytes(t)};
  return PTS_OK;
}

uint64_t pts_batch_storage_len(pts_batch* b, uint32_t spot) {
  return spot < b->spots.size() ? b->spots[spot]->tree.storage_len : 0;
}

pts_status pts_batch_get_storage(pts_batch* b, uint32_t spot, float* regrets, float* sums) {
  NEED_BATCH_ENGINE(b);
  if (spot >= b->spots.size()) return bad("invalid input: spot");
  return st(b->engine->read_storage(b->store_off[spot], b->spots[spot]->tree.storage_len, regrets, sums));
}

pts_status pts_batch_iterate(pts_batch* b, uint32_t count) {
  NEED_BATCH_ENGINE(b);
  for (uint32_t i = 0; i < count; i++) {
    b->iterations += 1;
    Status e = b->engine->cfr_iteration(discount(b->iterations, b->spots[0]->cfg));
    if (e != kOk) return st(e);
  }
  return PTS_OK;
}

pts_status pts_batch_exploitability(pts_batch* b, double* out) {
  NEED_BATCH_ENGINE(b);
  std::vector<double> x;
  pts_status r = batch_expl(b, std::vector<uint32_t>(b->spots.size(), 1u), x);
  if (r != PTS_OK) return r;
  std::memcpy(out, x.data(), x.size() * sizeof(double));
  return PTS_OK;
}

pts_status pts_batch_solve(pts_batch* b, uint32_t max_iterations, double target_pct, uint32_t check_every,
                           pts_progress_c* out) {
  // every spot on its own copy of the single solve's check schedule
  // (lib.rs:594-642): a check when the spot's next check is due, the gap
  // from its own last two checks (next_check_gap); only the spots due are
  // measured, and a spot at its target is frozen
  NEED_BATCH_ENGINE(b);
  const uint32_t K = (uint32_t)b->spots.size(), ce = std::max(check_every, 1u);
  for (uint32_t k = 0; k < K; k++)
    out[k] = {b->iterations, max_iterations, std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
              0, 0};
  std::vector<uint32_t> next(K, b->iterations + ce), lt(K, 0);
  std::vector<double> le(K, 0.0);
  std::vector<bool> has_last(K, false);
  uint32_t left = 0;
  for (uint32_t k = 0; k < K; k++) left += b->active[k];
  while (b->iterations < max_iterations && left > 0) {
    pts_status r = pts_batch_iterate(b, 1);
    if (r != PTS_OK) return r;
    const uint32_t t = b->iterations;
    const bool last = t >= max_iterations;
    std::vector<uint32_t> due(K, 0u);
    bool any = false;
    for (uint32_t k = 0; k < K; k++)
      if (b->active[k] && (t >= next[k] || last)) due[k] = 1u, any = true;
    if (!any) continue;
    std::vector<double> x;
    if ((r = batch_expl(b, due, x)) != PTS_OK) return r;
    for (uint32_t k = 0; k < K; k++) {
      if (!due[k]) continue;
      const double pot = (double)b->spots[k]->cfg.pot, pct = 100.0 * x[k] / pot;
      out[k].iteration = t;
      out[k].exploitability = x[k];
      out[k].exploitability_pct = pct;
      if (pct <= target_pct || last) {
        out[k].finished = 1;
        b->active[k] = 0;  // a solved spot is frozen: its traversals are skipped
        left--;
        continue;
      }
      next[k] = t + check_gap(has_last[k], lt[k], le[k], t, x[k], target_pct * pot / 100.0, ce);
      has_last[k] = true;
      lt[k] = t;
      le[k] = x[k];
    }
    Status e = b->engine->set_active(b->active);
    if (e != kOk) return st(e);
  }
  return PTS_OK;
}

// Times `iterations` DCFR iterations for every run length x down-kernel
// workgroup size, keeps the fastest, then restarts the solve from zero.
pts_status pts_batch_autotune(pts_batch* b, uint32_t iterations, uint32_t best[2], double* best_ms) {
  NEED_BATCH_ENGINE(b);
  static const uint32_t runs[] = {1, 2, 4, 8, 16, 32, 64, 256};
  static const uint32_t downs[] = {128, 256, 512, 1024};
  double bm = std::numeric_limits<double>::infinity();
  uint32_t br = b->engine->run_length(), bd = b->engine->down_threads();
  const uint32_t it = std::max(iterations, 1u);
  for (uint32_t rl : runs)
    for (uint32_t dt : downs) {
      Status e = b->engine->retune(rl, dt);
      if (e != kOk) return st(e);
      if ((e = b->engine->cfr_iteration(discount(1, b->spots[0]->cfg))) != kOk) return st(e);
      auto t0 = std::chrono::steady_clock::now();
      for (uint32_t i = 0; i < it; i++)
        if ((e = b->engine->cfr_iteration(discount(2 + i, b->spots[0]->cfg))) != kOk) return st(e);
      double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / it;
      if (ms < bm) {
        bm = ms;
        br = rl;
        bd = dt;
      }
    }
  Status e = b->engine->retune(br, bd);
  if (e != kOk) return st(e);
  if ((e = b->engine->zero_storage()) != kOk) return st(e);
  b->iterations = 0;
  b->active.assign(b->spots.size(), 1u);
  if ((e = b->engine->set_active(b->active)) != kOk) return st(e);
  best[0] = br;
  best[1] = bd;
  *best_ms = bm;
  return PTS_OK;
}

pts_status pts_batch_root_rows(pts_batch* b, int mode, double* out) {
  NEED_BATCH_ENGINE(b);
  if (mode != PTS_MODE_BEST_RESPONSE && mode != PTS_MODE_AVERAGE) return bad("invalid input: mode");
  const uint32_t K = (uint32_t)b->spots.size();
  std::vector<uint32_t> all(K, 1u);
  Status e;
  if ((e = b->engine->set_active(all)) != kOk) return st(e);
  std::vector<double> rows;
  if ((e = b->engine->root_rows((Mode)mode, rows)) != kOk) return st(e);
  if ((e = b->engine->set_active(b->active)) != kOk) return st(e);
  const uint32_t rl = b->engine->rowlen();
  for (uint32_t k = 0; k < K; k++)
    for (int p = 0; p < 2; p++) {
      double* o = out + ((size_t)k * 2 + p) * 1326;
      std::fill(o, o + 1326, 0.0);
      const Hands& h = b->spots[k]->hands;
      for (size_t i = 0; i < h.n(p); i++) o[h.combos[p][i]] = rows[((size_t)k * 2 + p) * rl + i];
    }
  return PTS_OK;
}

double pts_batch_num_combinations(pts_batch* b, uint32_t spot) {
  return spot < b->spots.size() ? b->spots[spot]->hands.num_combinations : 0.0;
}

pts_status pts_batch_set_tiles(pts_batch* b, uint32_t run_length, uint32_t down_threads) {
  NEED_BATCH_ENGINE(b);
  return st(b->engine->retune(run_length, down_threads));
}

pts_status pts_batch_profile_iteration(pts_batch* b, double ms[4], uint32_t launches[4], double* wall_ms) {
  NEED_BATCH_ENGINE(b);
  b->iterations += 1;
  Profile pr;
  Status e = b->engine->profile_iteration(discount(b->iterations, b->spots[0]->cfg), pr);
  if (e != kOk) return st(e);
  for (int i = 0; i < 4; i++) {
    ms[i] = pr.ms[i];
    launches[i] = pr.launches[i];
  }
  *wall_ms = pr.wall_ms;
  return PTS_OK;
}

pts_status pts_batch_host_info(pts_batch* b, uint64_t* num_nodes, uint64_t* storage_len, uint64_t scratch[2]) {
  const Plan& P = b->engine ? b->engine->plan() : b->host_plan;
  *num_nodes = P.num_nodes;
  *storage_len = P.storage_len;
  scratch[0] = b->host_scratch[0];
  scratch[1] = b->host_scratch[1];
  return PTS_OK;
}

// ------------------------------------------------------------------ value network
}  // extern "C"

struct pts_net {
  std::unique_ptr<net::Net> n;
};

extern "C" {

uint64_t pts_net_num_params(void) { return net::num_params(); }
double pts_net_forward_flops_per_row(void) { return net::forward_flops_per_row(); }

pts_status pts_net_create(int device, uint32_t batch_cap, uint64_t seed, pts_net** out) {
  *out = nullptr;
  std::unique_ptr<pts_net> h(new pts_net());
  Status e = net::Net::create(device, batch_cap, h->n);
  if (e != kOk) return st(e);
  std::vector<float> p;
  net::he_init(seed, p);
  if ((e = h->n->set_params(p.data())) != kOk) return st(e);
  *out = h.release();
  return PTS_OK;
}

void pts_net_destroy(pts_net* n) { delete n; }
pts_status pts_net_set_params(pts_net* n, const float* params) { return st(n->n->set_params(params)); }
pts_status pts_net_get_params(pts_net* n, float* params) { return st(n->n->get_params(params)); }
pts_status pts_net_get_grads(pts_net* n, float* grads) { return st(n->n->get_grads(grads)); }
pts_status pts_net_forward(pts_net* n, const float* x, uint32_t rows, float* y) { return st(n->n->forward(x, rows, y)); }
pts_status pts_net_forward_noalloc(pts_net* n, const float* x, uint32_t rows, float* y, void* stream) {
  return st(n->n->forward_noalloc(x, rows, y, (hipStream_t)stream));
}
pts_status pts_net_gradients(pts_net* n, const float* x, const float* target, const float* mask, uint32_t rows, double* loss) {
  return st(n->n->gradients(x, target, mask, rows, loss));
}
pts_status pts_net_train_step(pts_net* n, const float* x, const float* target, const float* mask, uint32_t rows, float lr,
                              double* loss) {
  return st(n->n->train_step(x, target, mask, rows, lr, loss));
}
pts_status pts_net_profile(pts_net* n, uint32_t rows, double* fwd_ms, double* train_ms) {
  return st(n->n->profile(rows, fwd_ms, train_ms));
}

// ------------------------------------------------------------------ depth-limited solve
// A flop (or turn) solver whose tree stops where the root street's betting
// ends with chips behind; the value network values those leaves per dealt
// card. The result is a pts_solver: iterate / solve / storage / nodes work
// as for a full tree (exploitability is that of the depth-limited game).
pts_status pts_dl_new(const pts_config_c* cfg, int device, int precision, pts_net* n, pts_solver** out) {
  *out = nullptr;
  std::unique_ptr<pts_solver> s(new pts_solver());
  Plan plan;
  std::string err = host_build(cfg, s.get(), plan, true);
  if (!err.empty()) return bad(err);
  if (s->compressed) return bad("invalid input: depth-limited solves need f32 storage");
  if (plan.leaves.empty()) return bad("invalid input: the depth-limited tree has no leaves");
  if (device < 0) {  // host only: the cut tree and its plan
    s->host_plan = std::move(plan);
    *out = s.release();
    return PTS_OK;
  }
  Status e = Engine::create(device, std::move(plan), nullptr, nullptr, precision, 0, s->engine);
  if (e != kOk) return st(e);
  if ((e = s->engine->attach_net(n ? n->n.get() : nullptr, s->spec.board)) != kOk) return st(e);
  *out = s.release();
  return PTS_OK;
}

uint32_t pts_solver_num_leaves(pts_solver* s) {
  const Plan& P = s->engine ? s->engine->plan() : s->host_plan;
  return (uint32_t)P.leaves.size();
}

// The root street's decision nodes of `src` (a depth-limited tree) copied
// into the same action paths of `dst` (the full tree): regrets times
// regret_scale and strategy sums times sum_scale. Returns the nodes copied.
pts_status pts_solver_copy_root_street(pts_solver* dst, pts_solver* src, float regret_scale, float sum_scale,
                                       uint32_t* copied) {
  std::scoped_lock lock_(dst->mu, src->mu);
  NEED_ENGINE(dst);
  NEED_ENGINE(src);
  if (dst->hands.n(0) != src->hands.n(0) || dst->hands.n(1) != src->hands.n(1))
    return bad("invalid input: the two solvers have different hand lists");
  const Tree& a = src->tree;
  const Tree& b = dst->tree;
  const uint8_t street = (uint8_t)src->spec.root_street;
  std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 0u}};
  uint32_t n = 0;
  while (!stack.empty()) {
    auto [x, y] = stack.back();
    stack.pop_back();
    const Node& na = a.nodes[x];
    const Node& nb = b.nodes[y];
    if (!is_player(na.kind) || na.kind != nb.kind || na.street != street || na.num_children != nb.num_children) continue;
    const size_t len = (size_t)na.num_children * src->hands.n(na.kind);
    std::vector<float> r(len), m(len);
    Status e = src->engine->read_storage(na.offset, len, r.data(), m.data());
    if (e != kOk) return st(e);
    for (auto& v : r) v *= regret_scale;
    for (auto& v : m) v *= sum_scale;
    if ((e = dst->engine->write_storage(nb.offset, len, r.data(), m.data())) != kOk) return st(e);
    n++;
    for (uint32_t k = 0; k < na.num_children; k++) stack.push_back({na.first_child + k, nb.first_child + k});
  }
  *copied = n;
  return PTS_OK;
}

// Pins dst's root-street decision nodes to src's average strategy: sums are
// copied, regrets become the average strategy's probabilities (the sums
// normalised per hand, uniform where they sum to zero, cfr.rs:124-154), so
// regret matching in dst plays exactly that strategy at those nodes.
pts_status pts_solver_pin_root_street(pts_solver* dst, pts_solver* src, uint32_t* pinned) {
  std::scoped_lock lock_(dst->mu, src->mu);
  NEED_ENGINE(dst);
  NEED_ENGINE(src);
  if (dst->hands.n(0) != src->hands.n(0) || dst->hands.n(1) != src->hands.n(1))
    return bad("invalid input: the two solvers have different hand lists");
  const Tree& a = src->tree;
  const Tree& b = dst->tree;
  const uint8_t street = (uint8_t)src->spec.root_street;
  std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 0u}};
  std::vector<uint32_t> frozen;
  uint32_t n = 0;
  while (!stack.empty()) {
    auto [x, y] = stack.back();
    stack.pop_back();
    const Node& na = a.nodes[x];
    const Node& nb = b.nodes[y];
    if (!is_player(na.kind) || na.kind != nb.kind || na.street != street || na.num_children != nb.num_children) continue;
    const size_t nh = src->hands.n(na.kind), na_ = na.num_children, len = na_ * nh;
    std::vector<float> r(len), m(len);
    Status e = src->engine->read_storage(na.offset, len, r.data(), m.data());
    if (e != kOk) return st(e);
    for (size_t h = 0; h < nh; h++) {
      double d = 0.0;
      for (size_t k = 0; k < na_; k++) d += m[k * nh + h];
      for (size_t k = 0; k < na_; k++) r[k * nh + h] = d > 0.0 ? (float)(m[k * nh + h] / d) : (float)(1.0 / (double)na_);
    }
    if ((e = dst->engine->write_storage(nb.offset, len, r.data(), m.data())) != kOk) return st(e);
    frozen.push_back(y);
    n++;
    for (uint32_t k = 0; k < na.num_children; k++) stack.push_back({na.first_child + k, nb.first_child + k});
  }
  // the pinned regrets stay as written: both traversals of every iteration
  // play the pinned strategy
  Status e = dst->engine->set_frozen(frozen, true);
  if (e != kOk) return st(e);
  *pinned = n;
  return PTS_OK;
}

}  // extern "C"
