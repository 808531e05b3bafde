// C API of the CPU oracle (test infrastructure only).
#include <omp.h>

#include <cstdio>
#include <cstring>
#include <set>

#include "oracle.h"
#include "oracle_core.hpp"

using namespace orc;

static void put_err(const std::string& e, char* err, int32_t errlen) {
  if (err && errlen > 0) {
    std::snprintf(err, (size_t)errlen, "%s", e.c_str());
  }
}

static StreetSizes sizes_from(const orc_sizes_c& bet, const orc_sizes_c& raise) {
  StreetSizes s;
  for (int i = 0; i < bet.count; i++) s.bet.push_back({(BetSize::Kind)bet.kind[i], bet.value[i]});
  for (int i = 0; i < raise.count; i++) s.raise.push_back({(BetSize::Kind)raise.kind[i], raise.value[i]});
  return s;
}

static PostflopMenu menu_from(const orc_menu_c& m) {
  PostflopMenu r;
  r.flop = sizes_from(m.bet[0], m.raise[0]);
  r.turn = sizes_from(m.bet[1], m.raise[1]);
  r.river = sizes_from(m.bet[2], m.raise[2]);
  r.max_raises = m.max_raises;
  r.add_allin_threshold = m.add_allin_threshold;
  r.force_allin_threshold = m.force_allin_threshold;
  r.merge_threshold = m.merge_threshold;
  return r;
}

static SolveConfig config_from(const orc_config_c* c) {
  SolveConfig s;
  s.board.assign(c->board, c->board + c->board_len);
  for (int p = 0; p < 2; p++) s.ranges[p].assign(c->ranges[p], c->ranges[p] + 1326);
  BettingState& b = s.root;
  b.street = (Street)c->street;
  b.pot = c->pot;
  b.commit[0] = c->commit[0]; b.commit[1] = c->commit[1];
  b.stack[0] = c->stack[0]; b.stack[1] = c->stack[1];
  b.current_bet = c->current_bet;
  b.last_raise_inc = c->last_raise_inc;
  b.num_raises = c->num_raises;
  b.to_act = c->to_act;
  b.needs[0] = c->needs[0] != 0; b.needs[1] = c->needs[1] != 0;
  s.menu = menu_from(c->menu);
  s.has_future_menu = c->has_future_menu != 0;
  if (s.has_future_menu) s.future_menu = menu_from(c->future_menu);
  s.big_blind = c->big_blind;
  s.prune = c->prune;
  s.max_memory_bytes = c->max_memory_bytes;
  s.rake_rate = c->rake_rate;
  s.rake_cap = c->rake_cap;
  s.storage_mode = c->storage_mode;
  s.isomorphism = c->isomorphism != 0;
  s.dcfr.alpha = c->alpha;
  s.dcfr.beta_some = c->beta_some != 0;
  s.dcfr.beta = c->beta;
  s.dcfr.gamma = c->gamma;
  s.dcfr.restart_average = c->restart_average != 0;
  s.dcfr.unweighted_average = c->unweighted_average != 0;
  s.backend = (Backend)c->backend;
  s.chance_mode = c->chance_mode;
  s.sample_turn = c->sample_turn;
  s.sample_river = c->sample_river;
  s.sample_seed = c->sample_seed;
  return s;
}

static void stats_to(const TreeStats& t, orc_stats_c* o) {
  o->player_nodes = t.player_nodes;
  o->chance_nodes = t.chance_nodes;
  o->terminal_nodes = t.terminal_nodes;
  o->storage_cells = t.storage_cells;
  o->memory_bytes = t.memory_bytes();
}

static std::vector<Step> path_from(const orc_step_c* path, int32_t n) {
  std::vector<Step> v;
  for (int i = 0; i < n; i++) {
    Step s;
    s.kind = (StepKind)path[i].kind;
    s.action = {(ActKind)path[i].act_kind, path[i].to};
    s.card = path[i].card;
    v.push_back(s);
  }
  return v;
}

extern "C" {

size_t orc_config_size(void) { return sizeof(orc_config_c); }

uint32_t orc_evaluate(const int32_t* cards, int32_t n) { return evaluate_cards(cards, n); }

void orc_evaluate_batch(const uint8_t* cards, int32_t ncards, uint64_t count, uint32_t* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)count; i++) {
    int c[7];
    for (int k = 0; k < ncards; k++) c[k] = cards[i * ncards + k];
    out[i] = evaluate_cards(c, ncards);
  }
}

uint64_t orc_enumerate_hands(int32_t ncards, uint64_t first, uint64_t count, uint8_t* out) {
  // unrank `first` (lexicographic order of ascending card tuples), then step
  auto binom = [](int n, int k) -> uint64_t {
    if (k < 0 || k > n) return 0;
    uint64_t r = 1;
    for (int i = 1; i <= k; i++) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
    return r;
  };
  const int k = ncards;
  int c[7];
  uint64_t rank = first;
  int lo = 0;
  for (int i = 0; i < k; i++) {
    int x = lo;
    for (;; x++) {
      const uint64_t below = binom(51 - x, k - i - 1);  // tuples with card x here
      if (x > 52 - (k - i)) return 0;                    // `first` beyond the last hand
      if (rank < below) break;
      rank -= below;
    }
    c[i] = x;
    lo = x + 1;
  }
  uint64_t n = 0;
  for (; n < count; n++) {
    for (int i = 0; i < k; i++) out[n * k + i] = (uint8_t)c[i];
    int i = k - 1;
    while (i >= 0 && c[i] == 52 - k + i) i--;
    if (i < 0) return n + 1;
    c[i]++;
    for (int j = i + 1; j < k; j++) c[j] = c[j - 1] + 1;
  }
  return n;
}
int32_t orc_combo_index(int32_t a, int32_t b) { return combo_index(a, b); }
void orc_combo_cards(int32_t idx, int32_t* hi, int32_t* lo) {
  *hi = combo_cards(idx).hi;
  *lo = combo_cards(idx).lo;
}
int32_t orc_combo_class(int32_t idx) { return combo_class(idx); }

int32_t orc_parse_range(const char* s, float* out) {
  std::vector<float> w;
  if (!parse_range(s, w)) return 1;
  std::memcpy(out, w.data(), 1326 * sizeof(float));
  return 0;
}

int32_t orc_parse_cards(const char* s, int32_t* out, int32_t cap) {
  std::vector<int> c = parse_cards(s);
  if (c.empty() && s[0]) return -1;
  for (size_t i = 0; i < c.size() && (int)i < cap; i++) out[i] = c[i];
  return (int32_t)c.size();
}

void orc_count_categories(int32_t ncards, uint64_t counts[9], uint64_t* distinct) {
  // Exhaustive enumeration of all n-card hands (n = 5 or 7), in parallel over
  // the first card.
  std::memset(counts, 0, 9 * sizeof(uint64_t));
  std::vector<std::vector<uint64_t>> local(omp_get_max_threads(), std::vector<uint64_t>(9, 0));
  std::vector<std::set<uint32_t>> vals(omp_get_max_threads());
#pragma omp parallel for schedule(dynamic, 1)
  for (int a = 0; a < 52; a++) {
    int tid = omp_get_thread_num();
    std::set<uint32_t>& vs = vals[tid];
    uint64_t ma = eval_mask_bit(a);
    for (int b = a + 1; b < 52; b++) {
      uint64_t mb = ma | eval_mask_bit(b);
      for (int c = b + 1; c < 52; c++) {
        uint64_t mc = mb | eval_mask_bit(c);
        for (int d = c + 1; d < 52; d++) {
          uint64_t md = mc | eval_mask_bit(d);
          for (int e = d + 1; e < 52; e++) {
            uint64_t me = md | eval_mask_bit(e);
            if (ncards == 5) {
              uint32_t v = evaluate_mask(me);
              local[tid][v >> 26]++;
              vs.insert(v);
              continue;
            }
            for (int f = e + 1; f < 52; f++) {
              uint64_t mf = me | eval_mask_bit(f);
              for (int g = f + 1; g < 52; g++) {
                uint32_t v = evaluate_mask(mf | eval_mask_bit(g));
                local[tid][v >> 26]++;
                vs.insert(v);
              }
            }
          }
        }
      }
    }
  }
  std::set<uint32_t> all;
  for (size_t t = 0; t < local.size(); t++) {
    for (int k = 0; k < 9; k++) counts[k] += local[t][k];
    all.insert(vals[t].begin(), vals[t].end());
  }
  *distinct = all.size();
}

int32_t orc_postflop_actions(int32_t street, int64_t pot, int64_t to_call, int64_t current_bet, int64_t max_to,
                             int64_t min_raise_to, uint32_t num_raises, const orc_menu_c* menu, int32_t* kinds,
                             int64_t* tos, int32_t cap) {
  Situation s{(Street)street, pot, to_call, current_bet, max_to, min_raise_to, num_raises, 0, 100};
  std::vector<Action> a = postflop_actions(s, menu_from(*menu));
  for (size_t i = 0; i < a.size() && (int)i < cap; i++) {
    kinds[i] = a[i].kind;
    tos[i] = a[i].to;
  }
  return (int32_t)a.size();
}

void orc_discount(uint32_t t, double alpha, int32_t beta_some, double beta, double gamma, int32_t restart,
                  int32_t unweighted, float out[3]) {
  DcfrParams p;
  p.alpha = alpha; p.beta_some = beta_some != 0; p.beta = beta; p.gamma = gamma;
  p.restart_average = restart != 0; p.unweighted_average = unweighted != 0;
  Discount d = make_discount(t, p);
  out[0] = d.pos; out[1] = d.neg; out[2] = d.strat;
}

uint32_t orc_next_check_gap(int32_t has_prev, uint32_t t0, double e0, uint32_t t, double e, double target, uint32_t max_gap) {
  return next_check_gap(has_prev != 0, t0, e0, t, e, target, max_gap);
}

void* orc_solver_new(const orc_config_c* cfg, char* err, int32_t errlen) {
  Solver* s = new Solver();
  std::string e = make_solver(config_from(cfg), *s);
  if (!e.empty()) {
    put_err(e, err, errlen);
    delete s;
    return nullptr;
  }
  return s;
}

int32_t orc_estimate(const orc_config_c* cfg, orc_stats_c* out, char* err, int32_t errlen) {
  TreeStats t;
  std::string e = estimate(config_from(cfg), t);
  if (!e.empty()) { put_err(e, err, errlen); return 1; }
  stats_to(t, out);
  return 0;
}

void orc_solver_free(void* s) { delete (Solver*)s; }
void orc_solver_set_parallel(void* s, int32_t on, int32_t threads) {
  ((Solver*)s)->parallel = on != 0;
  if (threads > 0) omp_set_num_threads(threads);
}
void orc_solver_set_batch_river(void* s, int32_t on) { ((Solver*)s)->batch_river = on != 0; }
void orc_solver_stats(void* s, orc_stats_c* out) { stats_to(((Solver*)s)->stats, out); }
uint64_t orc_solver_storage_len(void* s) { return ((Solver*)s)->tree.storage_len; }
uint64_t orc_solver_num_nodes(void* s) { return ((Solver*)s)->tree.nodes.size(); }
int32_t orc_solver_num_hands(void* s, int32_t p) { return (int32_t)((Solver*)s)->hands.len(p); }
double orc_solver_num_combinations(void* s) { return ((Solver*)s)->hands.num_combinations; }

void orc_solver_hands(void* s, int32_t p, uint16_t* combos, float* weights) {
  Solver* v = (Solver*)s;
  for (size_t i = 0; i < v->hands.len(p); i++) {
    combos[i] = v->hands.combos[p][i];
    weights[i] = v->hands.weights[p][i];
  }
}

// f32 storage as is; 16-bit storage decoded (storage.rs:209-246)
void orc_solver_get_storage(void* s, float* regrets, float* sums) {
  Solver* v = (Solver*)s;
  if (v->compressed()) {
    for (const Node& n : v->tree.nodes) {
      if (!is_player(n.kind)) continue;
      const size_t len = (size_t)n.num_children * v->hands.len(n.kind);
      const float rs = v->q.regret_scale[n.slot] / 32767.0f, ss = v->q.strat_scale[n.slot] / 65535.0f;
      for (size_t i = 0; i < len; i++) {
        regrets[n.offset + i] = (float)v->q.regrets[n.offset + i] * rs;
        sums[n.offset + i] = (float)v->q.strat[n.offset + i] * ss;
      }
    }
    return;
  }
  std::memcpy(regrets, v->regrets.data(), v->regrets.size() * sizeof(float));
  std::memcpy(sums, v->strat.data(), v->strat.size() * sizeof(float));
}
int32_t orc_solver_set_storage(void* s, const float* regrets, const float* sums) {
  Solver* v = (Solver*)s;
  if (v->compressed()) return -1;
  std::memcpy(v->regrets.data(), regrets, v->regrets.size() * sizeof(float));
  std::memcpy(v->strat.data(), sums, v->strat.size() * sizeof(float));
  return 0;
}
void orc_chacha_block(int32_t rounds, const uint32_t* in, uint32_t* out) { chacha_block(rounds, in, out); }
void orc_stdrng_u64(uint64_t seed, uint64_t n, uint64_t* out) {
  StdRng r(seed);
  for (uint64_t i = 0; i < n; i++) out[i] = r.next_u64();
}
void orc_pcg32_words(uint64_t state, uint64_t inc, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; i++) out[i] = pcg32_next(state, inc);
}
void orc_stream_words(const uint32_t* key, int32_t rounds, uint64_t n, uint32_t* out, uint64_t* out64) {
  StdRng a(key, rounds), b(key, rounds);
  for (uint64_t i = 0; i < n; i++) out[i] = a.next_u32();
  for (uint64_t i = 0; i < n / 2; i++) out64[i] = b.next_u64();
}
int32_t orc_sample_cards(const int32_t* available, int32_t n, int32_t k, uint64_t seed, int32_t* out) {
  StdRng r(seed);
  std::vector<int> av(available, available + n);
  std::vector<int> s = sample_cards(av, (size_t)k, r);
  for (size_t i = 0; i < s.size(); i++) out[i] = s[i];
  return (int32_t)s.size();
}
void orc_encode_block(int32_t sums, const float* v, uint64_t n, int16_t* rq, uint16_t* sq, float* scale) {
  encode_block(sums != 0, v, n, rq, sq, scale);
}
int32_t orc_solver_is_compressed(void* s) { return ((Solver*)s)->compressed() ? 1 : 0; }
// 16-bit storage: the stored integers and the per-player-node scales
int32_t orc_solver_get_storage_q(void* s, int16_t* regrets, uint16_t* sums, float* regret_scale, float* strat_scale) {
  Solver* v = (Solver*)s;
  if (!v->compressed()) return -1;
  std::memcpy(regrets, v->q.regrets.data(), v->q.regrets.size() * sizeof(int16_t));
  std::memcpy(sums, v->q.strat.data(), v->q.strat.size() * sizeof(uint16_t));
  std::memcpy(regret_scale, v->q.regret_scale.data(), v->q.regret_scale.size() * sizeof(float));
  std::memcpy(strat_scale, v->q.strat_scale.data(), v->q.strat_scale.size() * sizeof(float));
  return 0;
}
int32_t orc_solver_set_storage_q(void* s, const int16_t* regrets, const uint16_t* sums, const float* regret_scale,
                                 const float* strat_scale) {
  Solver* v = (Solver*)s;
  if (!v->compressed()) return -1;
  std::memcpy(v->q.regrets.data(), regrets, v->q.regrets.size() * sizeof(int16_t));
  std::memcpy(v->q.strat.data(), sums, v->q.strat.size() * sizeof(uint16_t));
  std::memcpy(v->q.regret_scale.data(), regret_scale, v->q.regret_scale.size() * sizeof(float));
  std::memcpy(v->q.strat_scale.data(), strat_scale, v->q.strat_scale.size() * sizeof(float));
  return 0;
}
uint32_t orc_solver_iterations(void* s) { return ((Solver*)s)->iterations; }
void orc_solver_set_iterations(void* s, uint32_t t) { ((Solver*)s)->iterations = t; }

void orc_solver_iterate(void* s, uint32_t count) {
  Solver* v = (Solver*)s;
  for (uint32_t i = 0; i < count; i++) {
    v->iterations += 1;
    v->cpu_iteration();
  }
}

double orc_solver_exploitability(void* s) { return ((Solver*)s)->exploitability(); }
double orc_solver_exploitability_f64(void* s) { return ((Solver*)s)->exploitability_f64(); }

void orc_solver_root_values(void* s, int32_t mode, int32_t f64, double out[2]) {
  Solver* v = (Solver*)s;
  if (f64) v->root_values_f64((ModeKind)mode, out);
  else v->root_values((ModeKind)mode, out);
}

void orc_solver_traverse_root(void* s, int32_t p, int32_t mode, uint32_t t, float* out) {
  Solver* v = (Solver*)s;
  Ctx c = v->ctx();
  Mode m{(ModeKind)mode, make_discount(t, v->config.dcfr)};
  traverse(c, 0, p, v->hands.weights[p ^ 1].data(), v->hands.len(p ^ 1), m, out, v->hands.len(p));
}

void orc_solver_traverse_root_f64(void* s, int32_t p, int32_t mode, double* out) {
  Solver* v = (Solver*)s;
  Ctx c = v->ctx();
  std::vector<double> opp(v->hands.weights[p ^ 1].begin(), v->hands.weights[p ^ 1].end());
  traverse_f64(c, 0, p, opp.data(), opp.size(), (ModeKind)mode, out, v->hands.len(p));
}

void orc_solver_solve(void* s, uint32_t max_iterations, double target_pct, uint32_t check_every, orc_progress_c* out,
                      orc_progress_c* history, int32_t history_cap, int32_t* history_len) {
  Solver* v = (Solver*)s;
  SolveOptions o{max_iterations, target_pct, check_every};
  std::vector<Progress> h;
  Progress p = v->solve(o, &h);
  out->iteration = p.iteration;
  out->max_iterations = p.max_iterations;
  out->exploitability = p.exploitability;
  out->exploitability_pct = p.exploitability_pct;
  out->finished = p.finished;
  if (history_len) *history_len = (int32_t)h.size();
  for (size_t i = 0; i < h.size() && (int)i < history_cap; i++) {
    history[i].iteration = h[i].iteration;
    history[i].max_iterations = h[i].max_iterations;
    history[i].exploitability = h[i].exploitability;
    history[i].exploitability_pct = h[i].exploitability_pct;
    history[i].finished = h[i].finished;
  }
}

void orc_solver_nodes(void* s, uint8_t* kind, uint8_t* street, uint16_t* num_children, uint32_t* first_child,
                      int32_t* invest, uint64_t* offset, uint32_t* slot, float* rake, float* chance_factor,
                      int32_t* edge_card, int32_t* edge_act_kind, int64_t* edge_to) {
  Solver* v = (Solver*)s;
  for (size_t i = 0; i < v->tree.nodes.size(); i++) {
    const Node& n = v->tree.nodes[i];
    kind[i] = n.kind;
    street[i] = (uint8_t)n.street;
    num_children[i] = n.num_children;
    first_child[i] = n.first_child;
    invest[2 * i] = n.invest[0];
    invest[2 * i + 1] = n.invest[1];
    offset[i] = n.offset;
    slot[i] = n.slot;
    rake[i] = n.rake;
    chance_factor[i] = n.chance_factor;
    edge_card[i] = n.edge_kind == 2 ? n.edge_card : -1;
    edge_act_kind[i] = n.edge_kind == 1 ? (int32_t)n.edge_action.kind : -1;
    edge_to[i] = n.edge_kind == 1 ? n.edge_action.to : 0;
  }
}

int32_t orc_solver_num_boards(void* s) { return (int32_t)((Solver*)s)->tree.boards.size(); }
void orc_solver_board(void* s, int32_t i, int32_t* cards5) {
  const std::vector<int>& b = ((Solver*)s)->tree.boards[i];
  for (size_t k = 0; k < 5; k++) cards5[k] = k < b.size() ? b[k] : -1;
}

int32_t orc_solver_query(void* s, const orc_step_c* path, int32_t npath, int32_t* player, int32_t* na, int32_t* n,
                         int32_t* act_kinds, int64_t* act_tos, uint16_t* combos, float* strategy, float* ev, float* reach,
                         double reach_fraction[2], int64_t* pot, int64_t invest[2], int32_t cap_actions, int32_t cap_hands,
                         char* err, int32_t errlen) {
  Solver* v = (Solver*)s;
  NodeInfo info;
  std::string e = v->query(path_from(path, npath), info);
  if (!e.empty()) { put_err(e, err, errlen); return 1; }
  *player = info.player;
  *na = (int32_t)info.actions.size();
  *n = (int32_t)info.combos.size();
  if (*na > cap_actions || *n > cap_hands) { put_err("capacity too small", err, errlen); return 2; }
  for (int a = 0; a < *na; a++) { act_kinds[a] = info.actions[a].kind; act_tos[a] = info.actions[a].to; }
  std::memcpy(combos, info.combos.data(), info.combos.size() * sizeof(uint16_t));
  std::memcpy(strategy, info.strategy.data(), info.strategy.size() * sizeof(float));
  std::memcpy(ev, info.ev.data(), info.ev.size() * sizeof(float));
  std::memcpy(reach, info.reach.data(), info.reach.size() * sizeof(float));
  reach_fraction[0] = info.reach_fraction[0];
  reach_fraction[1] = info.reach_fraction[1];
  *pot = info.pot;
  invest[0] = info.invest[0];
  invest[1] = info.invest[1];
  return 0;
}

int32_t orc_solver_ranges_after(void* s, const orc_step_c* path, int32_t npath, float* out0, float* out1, char* err,
                                int32_t errlen) {
  Solver* v = (Solver*)s;
  std::vector<float> out[2];
  std::string e = v->ranges_after(path_from(path, npath), out);
  if (!e.empty()) { put_err(e, err, errlen); return 1; }
  std::memcpy(out0, out[0].data(), 1326 * sizeof(float));
  std::memcpy(out1, out[1].data(), 1326 * sizeof(float));
  return 0;
}

struct HS { HandSet hs; };

void* orc_handset_new(const float* r0, const float* r1, const int32_t* board, int32_t nboard) {
  std::vector<float> ranges[2] = {std::vector<float>(r0, r0 + 1326), std::vector<float>(r1, r1 + 1326)};
  uint64_t b = 0;
  for (int i = 0; i < nboard; i++) b |= card_bit(board[i]);
  HS* h = new HS();
  h->hs = make_handset(ranges, b, 0.0f, nullptr);
  return h;
}
void orc_handset_free(void* h) { delete (HS*)h; }
int32_t orc_handset_len(void* h, int32_t p) { return (int32_t)((HS*)h)->hs.len(p); }
void orc_handset_cards(void* h, int32_t p, uint8_t* cards2) {
  const HandSet& hs = ((HS*)h)->hs;
  for (size_t i = 0; i < hs.len(p); i++) { cards2[2 * i] = hs.cards[p][i].first; cards2[2 * i + 1] = hs.cards[p][i].second; }
}
double orc_handset_num_combinations(void* h) { return ((HS*)h)->hs.num_combinations; }
void orc_handset_compatible_reach(void* h, int32_t p, const float* opp, double* out) {
  std::vector<double> v = ((HS*)h)->hs.compatible_reach(p, opp);
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}
void orc_eval_fold(void* h, int32_t p, const float* opp, double payoff, float* out) { eval_fold(((HS*)h)->hs, p, opp, payoff, out); }
void orc_eval_showdown(void* h, const int32_t* board5, int32_t p, const float* opp, double win, double lose, double tie,
                       float* out) {
  StrengthTable t = make_table(((HS*)h)->hs, std::vector<int>(board5, board5 + 5));
  eval_showdown(((HS*)h)->hs, t, p, opp, win, lose, tie, out);
}
void orc_fold_batch1(void* h, int32_t p, const float* opp, double payoff, float* out) {
  const float* r[1] = {opp};
  float* o[1] = {out};
  fold_batch(((HS*)h)->hs, p, 1, r, &payoff, o);
}
void orc_showdown_batch1(void* h, const int32_t* board5, int32_t p, const float* opp, double win, double lose, double tie,
                         int32_t has_fold, double fold_pay, float* out, float* fold_out) {
  StrengthTable t = make_table(((HS*)h)->hs, std::vector<int>(board5, board5 + 5));
  const float* r[1] = {opp};
  float* o[1] = {out};
  float* fo[1] = {fold_out};
  int hf[1] = {has_fold};
  std::vector<float> tmp(((HS*)h)->hs.len(p), 0.0f);
  for (size_t i = 0; i < tmp.size(); i++) out[i] = 0.0f;
  showdown_batch(((HS*)h)->hs, t, p, 1, r, &win, &lose, &tie, o, hf, &fold_pay, fo);
}

}  // extern "C"
