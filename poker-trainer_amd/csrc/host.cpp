// Host side: evaluator, bet menu, betting transitions, tree arena, hand sets.
#include "host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <tuple>
#include <unordered_map>

namespace pts {

// ----------------------------------------------------------------- sampled chance
void ChaChaRng::block(const uint32_t in[16], uint32_t out[16], int rounds) {
  uint32_t v[16];
  std::memcpy(v, in, sizeof v);
  auto rot = [](uint32_t x, int r) { return (x << r) | (x >> (32 - r)); };
  auto quarter = [&](int a, int b, int c, int d) {
    v[a] += v[b]; v[d] ^= v[a]; v[d] = rot(v[d], 16);
    v[c] += v[d]; v[b] ^= v[c]; v[b] = rot(v[b], 12);
    v[a] += v[b]; v[d] ^= v[a]; v[d] = rot(v[d], 8);
    v[c] += v[d]; v[b] ^= v[c]; v[b] = rot(v[b], 7);
  };
  for (int r = 0; r < rounds; r += 2) {
    for (int col = 0; col < 4; col++) quarter(col, col + 4, col + 8, col + 12);
    for (int dg = 0; dg < 4; dg++) quarter(dg, 4 + (dg + 1) % 4, 8 + (dg + 2) % 4, 12 + (dg + 3) % 4);
  }
  for (int w = 0; w < 16; w++) out[w] = v[w] + in[w];
}

ChaChaRng::ChaChaRng(uint64_t seed) {
  uint64_t st = seed;
  for (uint32_t& k : key_) {
    st = st * 6364136223846793005ull + 11634580027462260723ull;
    const uint32_t x = (uint32_t)(((st >> 18) ^ st) >> 27);
    const unsigned r = (unsigned)(st >> 59);
    k = r ? (x >> r) | (x << (32 - r)) : x;
  }
}

uint32_t ChaChaRng::next_word() {
  if (pos_ == 16) {
    const uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                             key_[0], key_[1], key_[2], key_[3], key_[4], key_[5], key_[6], key_[7],
                             (uint32_t)blocks_, (uint32_t)(blocks_ >> 32), 0u, 0u};
    block(in, words_, 12);
    blocks_++;
    pos_ = 0;
  }
  return words_[pos_++];
}

// The stream is one sequence of words whatever the buffering, so two
// consecutive words make a u64 (rand_core's BlockRng reads the same pair,
// also across its 4-block buffer's end).
uint64_t ChaChaRng::next_u64() {
  const uint64_t lo = next_word();
  return (uint64_t)next_word() << 32 | lo;
}

std::vector<int> stratified_cards(const std::vector<int>& available, size_t k, ChaChaRng& rng) {
  if (k >= available.size()) return available;
  std::vector<int> suit[4];
  for (int c : available) suit[c & 3].push_back(c);
  for (auto& v : suit) std::sort(v.begin(), v.end());  // by rank: one suit's cards differ in rank only
  const double n = (double)available.size();
  size_t quota[4];
  std::tuple<double, double, int> rem[4];
  size_t given = 0;
  for (int s = 0; s < 4; s++) {
    const double share = (double)k * (double)suit[s].size() / n;
    const double whole = std::floor(share);
    quota[s] = (size_t)whole;
    given += quota[s];
    rem[s] = std::make_tuple(share - whole, rng.next_f64(), s);
  }
  std::sort(rem, rem + 4, [](const auto& a, const auto& b) { return a > b; });
  for (const auto& r : rem) {
    if (given >= k) break;
    const int s = std::get<2>(r);
    if (quota[s] < suit[s].size()) {
      quota[s]++;
      given++;
    }
  }
  std::vector<int> out;
  for (int s = 0; s < 4; s++) {
    if (quota[s] == 0) continue;
    const size_t m = suit[s].size();
    const double step = (double)m / (double)quota[s];
    const double start = rng.next_f64() * step;
    for (size_t j = 0; j < quota[s]; j++) out.push_back(suit[s][std::min((size_t)std::floor(start + (double)j * step), m - 1)]);
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

void combo_pair(int combo, int& hi, int& lo) {
  int h = 1;
  while ((h + 1) * h / 2 <= combo) h++;
  hi = h;
  lo = combo - h * (h - 1) / 2;
}

// ----------------------------------------------------------------- evaluator
// Table-free classification from the four per-suit rank lanes (the same
// encoding as crates/holdem/src/eval.rs:142-198: category << 26 | payload).
static inline uint32_t msb(uint32_t x) { return 31u - (uint32_t)__builtin_clz(x); }
static inline uint32_t keep_top(uint32_t x, int k) {
  uint32_t out = 0;
  while (k-- > 0 && x) { uint32_t b = 1u << msb(x); out |= b; x &= ~b; }
  return out;
}
static inline int straight_top(uint32_t r) {
  uint32_t run = r & (r << 1) & (r << 2) & (r << 3) & (r << 4);
  if (run) return (int)msb(run);
  return (r & 0x100Fu) == 0x100Fu ? 3 : -1;  // A-5 wheel
}

uint32_t hand_strength(uint64_t m) {
  uint32_t s0 = (uint32_t)(m & 0x1FFF), s1 = (uint32_t)(m >> 16 & 0x1FFF), s2 = (uint32_t)(m >> 32 & 0x1FFF),
           s3 = (uint32_t)(m >> 48 & 0x1FFF);
  const uint32_t lanes[4] = {s0, s1, s2, s3};
  for (uint32_t s : lanes)
    if (__builtin_popcount(s) >= 5) {
      int h = straight_top(s);
      return h >= 0 ? (8u << 26) | (uint32_t)h : (5u << 26) | keep_top(s, 5);
    }
  uint32_t any = s0 | s1 | s2 | s3;
  uint32_t pair = (s0 & s1) | (s0 & s2) | (s0 & s3) | (s1 & s2) | (s1 & s3) | (s2 & s3);
  uint32_t trip = (s0 & s1 & s2) | (s0 & s1 & s3) | (s0 & s2 & s3) | (s1 & s2 & s3);
  uint32_t quad = s0 & s1 & s2 & s3;
  if (quad) { uint32_t q = msb(quad); return (7u << 26) | q << 13 | keep_top(any & ~(1u << q), 1); }
  if (trip) {
    uint32_t t = msb(trip), rest = pair & ~(1u << t);
    if (rest) return (6u << 26) | t << 13 | 1u << msb(rest);
  }
  int st = straight_top(any);
  if (st >= 0) return (4u << 26) | (uint32_t)st;
  if (trip) { uint32_t t = msb(trip); return (3u << 26) | t << 13 | keep_top(any & ~(1u << t), 2); }
  if (__builtin_popcount(pair) >= 2) { uint32_t pp = keep_top(pair, 2); return (2u << 26) | pp << 13 | keep_top(any & ~pp, 1); }
  if (pair) { uint32_t p = msb(pair); return (1u << 26) | p << 13 | keep_top(any & ~(1u << p), 3); }
  return keep_top(any, 5);
}

// ----------------------------------------------------------------- menu
static inline Chips round_half_away(double x) { return (Chips)std::round(x); }

std::vector<Action> menu_actions(const Betting& b, const Menu& m, Chips bb) {
  int p = b.to_act, o = p ^ 1;
  // situation (crates/solver/src/tree.rs:63-77)
  Chips to_call = std::min(std::max(b.current_bet - b.commit[p], (Chips)0), b.stack[p]);
  Chips max_to = std::min(b.commit[p] + b.stack[p], b.commit[o] + b.stack[o]);
  Chips min_raise_to = b.current_bet + std::max(b.last_raise_inc, bb);
  std::vector<Action> out;
  if (to_call > 0) { out.push_back({kFold, 0}); out.push_back({kCall, 0}); }
  else out.push_back({kCheck, 0});
  // menu.rs:272-307
  if (!(max_to > b.current_bet) || b.num_raises >= m.max_raises) return out;
  int si = b.street == kTurn ? 1 : b.street == kRiver ? 2 : 0;
  const std::vector<Size>& list = b.current_bet == 0 ? m.bet[si] : m.raise[si];
  double pac = (double)(b.pot + to_call);
  std::vector<std::pair<Chips, size_t>> cand;
  for (size_t i = 0; i < list.size(); i++) {
    const Size& z = list[i];
    if (z.kind == 0) cand.push_back({b.current_bet + round_half_away(z.v * pac), i});
    else if (z.kind == 1) { if (b.current_bet != 0) cand.push_back({round_half_away(z.v * (double)b.current_bet), i}); }
    else cand.push_back({max_to, i});
  }
  if (m.add_allin > 0.0 && (double)(max_to - b.current_bet) <= m.add_allin * pac) cand.push_back({max_to, list.size()});
  // menu.rs:355-392: clamp, force all-in, drop non-raises, sort + dedup, merge
  Chips lo = std::min(min_raise_to, max_to);
  Chips pot_after = b.pot + to_call;
  std::vector<std::pair<Chips, size_t>> v;
  for (auto c : cand) {
    Chips a = std::min(std::max(c.first, lo), max_to);
    Chips thr = round_half_away((double)(pot_after + 2 * (a - b.current_bet)) * m.force_allin);
    if (max_to - a <= thr) a = max_to;
    if (a > b.current_bet) v.push_back({a, c.second});
  }
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end(), [](auto& x, auto& y) { return x.first == y.first; }), v.end());
  if (m.merge > 0.0 && v.size() > 1) {
    auto ratio = [&](Chips a) { return (double)(a - b.current_bet) / (double)pot_after; };
    std::vector<std::pair<Chips, size_t>> kept;
    for (size_t i = v.size(); i-- > 0;)
      if (kept.empty() || !((1.0 + ratio(kept.back().first)) / (1.0 + ratio(v[i].first)) < 1.0 + m.merge)) kept.push_back(v[i]);
    std::reverse(kept.begin(), kept.end());
    v.swap(kept);
  }
  for (auto& x : v) out.push_back({kRaise, x.first});
  return out;
}

int betting_apply(const Betting& b, const Action& a, Betting& s, Chips& put) {  // tree.rs:93-125
  int p = b.to_act, o = p ^ 1;
  s = b;
  put = 0;
  if (a.kind == kFold) return 1;
  if (a.kind == kCall) put = std::min(std::max(s.current_bet - s.commit[p], (Chips)0), s.stack[p]);
  if (a.kind == kRaise) {
    put = a.to - s.commit[p];
    if (a.to - s.current_bet >= s.last_raise_inc) s.last_raise_inc = a.to - s.current_bet;
    s.current_bet = a.to;
    s.num_raises++;
    s.needs[o] = s.stack[o] > 0;
  }
  s.commit[p] += put;
  s.stack[p] -= put;
  s.pot += put;
  s.needs[p] = false;
  if (s.needs[o]) { s.to_act = o; return 0; }
  return 2;
}

// ----------------------------------------------------------------- tree
namespace {
struct TreeWalker {
  const TreeSpec& spec;
  Tree* tree;  // null: count only
  TreeCounts counts;
  std::vector<int> dealt;
  std::unordered_map<uint64_t, uint32_t> board_of;
  ChaChaRng rng;  // sampled chance (tree.rs:415-429)

  const Menu& menu_for(Street s) const { return s == spec.root_street ? spec.menu : spec.future_menu; }
  float rake(Chips pot) const { return (float)std::min((double)pot * spec.rake_rate, spec.rake_cap); }
  Node blank() const {
    Node n{};
    n.kind = kShowdown; n.street = kRiver; n.chance_factor = 1.0f; n.card = -1; n.action = {kCheck, 0};
    return n;
  }
  size_t alloc(size_t k) {
    size_t f = tree->nodes.size();
    tree->nodes.resize(f + k, blank());
    return f;
  }

  void player(size_t id, const Betting& st, const int32_t inv[2]) {
    std::vector<Action> acts = menu_actions(st, menu_for(st.street), spec.big_blind);
    counts.player++;
    counts.cells += acts.size() * spec.hands[st.to_act];
    size_t first = 0;
    if (tree) {
      first = alloc(acts.size());
      Node& n = tree->nodes[id];
      n.kind = (uint8_t)st.to_act;
      n.street = st.street;
      n.num_children = (uint16_t)acts.size();
      n.first_child = (uint32_t)first;
      n.invest[0] = inv[0]; n.invest[1] = inv[1];
      n.offset = tree->storage_len;
      n.pslot = tree->player_nodes;  // tree.rs:463-469
      tree->storage_len += acts.size() * spec.hands[st.to_act];
      tree->player_nodes++;
    }
    for (size_t i = 0; i < acts.size(); i++) {
      Betting nx;
      Chips put;
      int tr = betting_apply(st, acts[i], nx, put);
      int32_t ci[2] = {inv[0], inv[1]};
      ci[st.to_act] += (int32_t)put;
      size_t child = first + i;
      if (tree) tree->nodes[child].action = acts[i];
      if (tr == 0) {
        player(child, nx, ci);
      } else if (tr == 1) {
        counts.terminal++;
        if (tree) {
          int f = st.to_act;
          Chips uncalled = std::max(st.commit[f ^ 1] - st.commit[f], (Chips)0);
          Node& n = tree->nodes[child];
          n.kind = (uint8_t)(kFold0 + f);
          n.street = st.street;
          n.invest[0] = ci[0]; n.invest[1] = ci[1];
          n.rake = rake(st.pot - uncalled);
        }
      } else {
        street_over(child, nx, ci);
      }
    }
  }

  void street_over(size_t id, const Betting& st, const int32_t inv[2]) {
    if (st.street == kRiver) {
      counts.terminal++;
      if (!tree) return;
      uint64_t key = 0;
      for (int c : dealt) key |= 1ull << c;
      auto it = board_of.find(key);
      uint32_t b;
      if (it == board_of.end()) {
        b = (uint32_t)tree->boards.size();
        std::vector<int> full = spec.board;
        full.insert(full.end(), dealt.begin(), dealt.end());
        tree->boards.push_back(full);
        board_of.emplace(key, b);
      } else {
        b = it->second;
      }
      Node& n = tree->nodes[id];
      n.kind = kShowdown; n.street = kRiver; n.invest[0] = inv[0]; n.invest[1] = inv[1];
      n.offset = b; n.rake = rake(st.pot);
      return;
    }
    bool all_in = st.stack[0] == 0 || st.stack[1] == 0;
    if (spec.depth_limit && !all_in && st.street == spec.root_street) {
      counts.terminal++;
      if (!tree) return;
      size_t board_len = spec.board.size() + dealt.size();
      float cards = (float)(52 - board_len), compatible = (float)(52 - board_len - 4);
      Node& n = tree->nodes[id];
      n.kind = kLeaf; n.street = st.street; n.invest[0] = inv[0]; n.invest[1] = inv[1];
      n.offset = (uint64_t)st.pot;
      n.chance_factor = 1.0f / (cards * compatible / cards);
      return;
    }
    counts.chance++;
    Street next = (Street)(st.street + 1);
    uint64_t used = 0;
    for (int c : spec.board) used |= 1ull << c;
    for (int c : dealt) used |= 1ull << c;
    std::vector<int> avail, cards;
    for (int c = 0; c < 52; c++) if (!(used >> c & 1)) avail.push_back(c);
    const size_t k = next == kTurn ? spec.sample_turn : spec.sample_river;
    const bool sample = spec.sampled && !all_in && k < avail.size();  // tree.rs:326-333
    // iso.rs:277-300: a card whose suit has a lower symmetric partner with
    // the same board ranks is merged onto that partner's card
    std::vector<IsoRef> refs;
    std::vector<int> ref_canon;
    uint16_t m[4] = {0, 0, 0, 0};
    for (int c : spec.board) m[c & 3] |= (uint16_t)(1u << (c >> 2));
    for (int c : dealt) m[c & 3] |= (uint16_t)(1u << (c >> 2));
    const bool iso = !sample && spec.iso && spec.iso->enabled && spec.iso->any();
    if (sample) cards = stratified_cards(avail, k, rng);
    for (int c : sample ? std::vector<int>() : avail) {
      int found = -1, pair = -1;
      if (iso)
        for (int s0 = 0; s0 < (c & 3) && found < 0; s0++)
          for (int pi = 0; pi < 6; pi++)
            if (kSuitPairs[pi][0] == s0 && kSuitPairs[pi][1] == (c & 3) && spec.iso->symmetric[pi] && m[s0] == m[c & 3]) {
              found = s0;
              pair = pi;
            }
      if (found >= 0) {
        refs.push_back({0, (uint8_t)pair, (int8_t)c});
        ref_canon.push_back((c >> 2) * 4 + found);
      } else {
        cards.push_back(c);
      }
    }
    size_t first = 0;
    if (tree) {
      first = alloc(cards.size());
      size_t board_len = spec.board.size() + dealt.size();
      // tree.rs:519-542: 1 / (cards * compatible / available), in f32
      float compatible = (float)(52 - board_len - 4);
      Node& n = tree->nodes[id];
      n.kind = kChance; n.street = st.street; n.num_children = (uint16_t)cards.size(); n.first_child = (uint32_t)first;
      n.invest[0] = inv[0]; n.invest[1] = inv[1];
      n.chance_factor = 1.0f / ((float)(cards.size() + refs.size()) * compatible / (float)avail.size());
      n.iso_first = (uint32_t)tree->iso_refs.size();
      n.iso_cnt = (uint16_t)refs.size();
      for (size_t r = 0; r < refs.size(); r++) {
        refs[r].child = (uint16_t)(std::find(cards.begin(), cards.end(), ref_canon[r]) - cards.begin());
        tree->iso_refs.push_back(refs[r]);
      }
    }
    for (size_t i = 0; i < cards.size(); i++) {
      size_t child = first + i;
      if (tree) tree->nodes[child].card = (int8_t)cards[i];
      dealt.push_back(cards[i]);
      if (all_in) {
        Betting s2 = st;
        s2.street = next;
        street_over(child, s2, inv);
      } else {
        Betting s2{};
        s2.street = next; s2.pot = st.pot; s2.stack[0] = st.stack[0]; s2.stack[1] = st.stack[1];
        s2.last_raise_inc = spec.big_blind; s2.to_act = 0;
        s2.needs[0] = st.stack[0] > 0; s2.needs[1] = st.stack[1] > 0;
        player(child, s2, inv);
      }
      dealt.pop_back();
    }
  }
};
}  // namespace

TreeCounts count_tree(const TreeSpec& spec) {
  TreeWalker w{spec, nullptr, {}, {}, {}, ChaChaRng(0)};  // tree.rs:337-343
  int32_t inv[2] = {0, 0};
  w.player(0, spec.root, inv);
  return w.counts;
}

Tree build_tree(const TreeSpec& spec) {
  Tree t;
  TreeWalker w{spec, &t, {}, {}, {}, ChaChaRng(spec.sampled ? spec.sample_seed : 0)};
  t.nodes.push_back(w.blank());
  int32_t inv[2] = {0, 0};
  w.player(0, spec.root, inv);
  return t;
}

// ----------------------------------------------------------------- isomorphism
Iso make_iso(const Hands& h, bool enabled) {  // iso.rs:245-273
  Iso iso;
  iso.enabled = enabled;
  if (!enabled) return iso;
  for (int pi = 0; pi < 6; pi++) {
    const int a = kSuitPairs[pi][0], b = kSuitPairs[pi][1];
    bool ok = true;
    std::vector<uint16_t> pp[2];
    for (int p = 0; p < 2 && ok; p++) {
      float mx = 0.0f;
      for (float w : h.weights[p]) mx = std::max(mx, w);
      mx = std::max(mx, 1e-30f);
      for (size_t i = 0; i < h.n(p) && ok; i++) {
        int c[2];
        for (int k = 0; k < 2; k++) {
          const int x = h.cards[p][2 * i + k], s = x & 3;
          c[k] = (x >> 2) * 4 + (s == a ? b : s == b ? a : s);
        }
        const uint32_t j = h.lookup[p][combo_of(c[0], c[1])];
        if (j == kNone || std::fabs(h.weights[p][i] - h.weights[p][j]) > 1e-4f * mx) ok = false;
        else pp[p].push_back((uint16_t)j);
      }
    }
    if (!ok) continue;
    iso.symmetric[pi] = true;
    iso.perm[0][pi] = std::move(pp[0]);
    iso.perm[1][pi] = std::move(pp[1]);
  }
  return iso;
}

// ----------------------------------------------------------------- hands
Hands make_hands(const float* const ranges[2], uint64_t board) {
  Hands h;
  for (int p = 0; p < 2; p++) {
    h.lookup[p].assign(1326, kNone);
    for (int c = 0; c < 1326; c++) {
      int a, b;
      combo_pair(c, a, b);
      if ((board >> a & 1) || (board >> b & 1)) continue;
      float w = ranges[p][c];
      if (!(w > 0.0f)) continue;
      h.lookup[p][c] = (uint32_t)h.combos[p].size();
      h.combos[p].push_back((uint16_t)c);
      h.cards[p].push_back((uint8_t)a);
      h.cards[p].push_back((uint8_t)b);
      h.weights[p].push_back(w);
    }
  }
  for (int p = 0; p < 2; p++)
    for (uint16_t c : h.combos[p]) h.same[p].push_back(h.lookup[p ^ 1][c]);
  std::vector<double> cr = compatible_reach(h, 0, h.weights[1].data());
  double s = 0.0;
  for (size_t i = 0; i < cr.size(); i++) s += cr[i] * (double)h.weights[0][i];  // hands.rs:95-98
  h.num_combinations = s;
  return h;
}

std::vector<double> compatible_reach(const Hands& h, int p, const float* opp) {  // hands.rs:102-125
  int o = p ^ 1;
  double total = 0.0, minus[52] = {0};
  for (size_t k = 0; k < h.n(o); k++) {
    if (opp[k] == 0.0f) continue;
    double w = (double)opp[k];
    total += w;
    minus[h.cards[o][2 * k]] += w;
    minus[h.cards[o][2 * k + 1]] += w;
  }
  std::vector<double> out(h.n(p));
  for (size_t i = 0; i < h.n(p); i++) {
    double s = h.same[p][i] == kNone ? 0.0 : (double)opp[h.same[p][i]];
    out[i] = std::max(total - minus[h.cards[p][2 * i]] - minus[h.cards[p][2 * i + 1]] + s, 0.0);
  }
  return out;
}

}  // namespace pts
