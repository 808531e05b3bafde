// CPU oracle: C++ restatement of the reference CPU solve (test infrastructure
// only; see oracle_core.hpp for the scope, the pins and what stays unpinned).
#include "oracle_core.hpp"

#include <cctype>
#include <functional>
#include <memory>
#include <omp.h>

namespace orc {

// =================================================================== cards
static int parse_rank_ch(char c) {  // card.rs:103-106
  static const char R[] = "23456789TJQKA";
  c = (char)std::toupper((unsigned char)c);
  for (int i = 0; i < 13; i++) if (R[i] == c) return i;
  return -1;
}
static int parse_suit_ch(char c) {  // card.rs:108-111
  static const char S[] = "cdhs";
  c = (char)std::tolower((unsigned char)c);
  for (int i = 0; i < 4; i++) if (S[i] == c) return i;
  return -1;
}

int parse_card(const std::string& s0) {  // card.rs:113-126
  std::string s;
  for (char c : s0) if (!std::isspace((unsigned char)c)) s.push_back(c);
  if (s.size() != 2) return -1;
  int r = parse_rank_ch(s[0]), u = parse_suit_ch(s[1]);
  if (r < 0 || u < 0) return -1;
  return r * 4 + u;
}

std::vector<int> parse_cards(const std::string& s) {  // card.rs:129-136
  std::string c;
  for (char ch : s) if (!std::isspace((unsigned char)ch) && ch != ',') c.push_back(ch);
  std::vector<int> out;
  if (c.size() % 2) return {};
  for (size_t i = 0; i < c.size(); i += 2) {
    int k = parse_card(c.substr(i, 2));
    if (k < 0) return {};
    out.push_back(k);
  }
  return out;
}

// =================================================================== combos
static std::vector<ComboCards> build_combos() {  // combo.rs:19-41
  std::vector<ComboCards> v(1326);
  for (int hi = 1; hi < 52; hi++)
    for (int lo = 0; lo < hi; lo++) v[hi * (hi - 1) / 2 + lo] = {(uint8_t)hi, (uint8_t)lo};
  return v;
}
static const std::vector<ComboCards>& combos_table() {
  static const std::vector<ComboCards> t = build_combos();
  return t;
}
const ComboCards& combo_cards(int idx) { return combos_table()[idx]; }

int combo_class(int i) {  // combo.rs:464-485
  const ComboCards& c = combo_cards(i);
  int hi = rank_of(c.hi), lo = rank_of(c.lo);
  bool suited = suit_of(c.hi) == suit_of(c.lo);
  int rh = 12 - hi, rl = 12 - lo;
  if (hi == lo) return rh * 13 + rh;
  if (suited) return rh * 13 + rl;
  return rl * 13 + rh;
}

HandClass HandClass::make(int r1, int r2, bool suited) {  // combo.rs:69-73
  int hi = r1 >= r2 ? r1 : r2, lo = r1 >= r2 ? r2 : r1;
  return {hi, lo, suited && hi != lo};
}
int HandClass::grid() const {  // combo.rs:90-97
  int rh = 12 - hi, rl = 12 - lo;
  return suited ? rh * 13 + rl : rl * 13 + rh;
}
HandClass HandClass::from_grid(int g) {  // combo.rs:99-109
  int row = g / 13, col = g % 13;
  int r1 = 12 - row, r2 = 12 - col;
  if (row == col) return make(r1, r2, false);
  if (row < col) return make(r1, r2, true);
  return make(r1, r2, false);
}
std::vector<int> HandClass::combos() const {  // combo.rs:127-149
  std::vector<int> out;
  for (int s1 = 0; s1 < 4; s1++)
    for (int s2 = 0; s2 < 4; s2++) {
      int a = hi * 4 + s1, b = lo * 4 + s2;
      if (a == b) continue;
      bool ok = (hi == lo) ? s1 < s2 : suited ? s1 == s2 : s1 != s2;
      if (ok) out.push_back(combo_index(a, b));
    }
  return out;
}

// =================================================================== evaluator
uint64_t eval_mask_bit(int c) { return 1ull << (suit_of(c) * 16 + rank_of(c)); }  // eval.rs:75-78

static inline uint32_t high_bit(uint32_t x) { return 31 - __builtin_clz(x); }
static inline uint32_t top_bits(uint32_t x, int n) {  // eval.rs:108-121
  uint32_t out = 0;
  for (int i = 0; i < n; i++) {
    if (x == 0) break;
    uint32_t b = 1u << high_bit(x);
    out |= b;
    x ^= b;
  }
  return out;
}
static inline int straight_high(uint32_t ranks) {  // eval.rs:123-135; -1 = none
  const uint32_t WHEEL = 0x100F;
  uint32_t run = ranks & (ranks << 1) & (ranks << 2) & (ranks << 3) & (ranks << 4);
  if (run) return (int)high_bit(run);
  if ((ranks & WHEEL) == WHEEL) return 3;
  return -1;
}
static inline uint32_t hv(uint32_t cat, uint32_t payload) { return (cat << 26) | payload; }

uint32_t evaluate_mask(uint64_t mask) {  // eval.rs:142-198
  uint32_t s[4] = {(uint32_t)(mask & 0x1FFF), (uint32_t)((mask >> 16) & 0x1FFF),
                   (uint32_t)((mask >> 32) & 0x1FFF), (uint32_t)((mask >> 48) & 0x1FFF)};
  for (int i = 0; i < 4; i++) {
    if (__builtin_popcount(s[i]) >= 5) {
      int h = straight_high(s[i]);
      if (h >= 0) return hv(8, (uint32_t)h);
      return hv(5, top_bits(s[i], 5));
    }
  }
  uint32_t any = s[0] | s[1] | s[2] | s[3];
  uint32_t two = (s[0] & s[1]) | (s[0] & s[2]) | (s[0] & s[3]) | (s[1] & s[2]) | (s[1] & s[3]) | (s[2] & s[3]);
  uint32_t three = (s[0] & s[1] & s[2]) | (s[0] & s[1] & s[3]) | (s[0] & s[2] & s[3]) | (s[1] & s[2] & s[3]);
  uint32_t four = s[0] & s[1] & s[2] & s[3];
  if (four) {
    uint32_t q = high_bit(four);
    return hv(7, (q << 13) | top_bits(any & ~(1u << q), 1));
  }
  if (three) {
    uint32_t t = high_bit(three);
    uint32_t rest = two & ~(1u << t);
    if (rest) return hv(6, (t << 13) | (1u << high_bit(rest)));
  }
  int sh = straight_high(any);
  if (sh >= 0) return hv(4, (uint32_t)sh);
  if (three) {
    uint32_t t = high_bit(three);
    return hv(3, (t << 13) | top_bits(any & ~(1u << t), 2));
  }
  if (__builtin_popcount(two) >= 2) {
    uint32_t pairs = top_bits(two, 2);
    return hv(2, (pairs << 13) | top_bits(any & ~pairs, 1));
  }
  if (two) {
    uint32_t p = high_bit(two);
    return hv(1, (p << 13) | top_bits(any & ~(1u << p), 3));
  }
  return hv(0, top_bits(any, 5));
}

uint32_t evaluate_cards(const int* cards, int n) {
  uint64_t m = 0;
  for (int i = 0; i < n; i++) m |= eval_mask_bit(cards[i]);
  return evaluate_mask(m);
}

// =================================================================== ranges
namespace {
struct Pattern { int hi, lo; int suited; };  // suited: -1 none, 0 offsuit, 1 suited (range.rs:154-199)
bool pattern_parse(const std::string& s, Pattern& p) {
  if (s.size() < 2 || s.size() > 3) return false;
  int r1 = parse_rank_ch(s[0]), r2 = parse_rank_ch(s[1]);
  if (r1 < 0 || r2 < 0) return false;
  p.hi = r1 >= r2 ? r1 : r2;
  p.lo = r1 >= r2 ? r2 : r1;
  if (s.size() == 2) { p.suited = -1; return true; }
  char c = (char)std::tolower((unsigned char)s[2]);
  if (c == 's' && p.hi != p.lo) { p.suited = 1; return true; }
  if (c == 'o' && p.hi != p.lo) { p.suited = 0; return true; }
  return false;
}
std::vector<HandClass> pattern_classes(const Pattern& p) {
  if (p.hi == p.lo) return {HandClass::make(p.hi, p.lo, false)};
  if (p.suited >= 0) return {HandClass::make(p.hi, p.lo, p.suited == 1)};
  return {HandClass::make(p.hi, p.lo, true), HandClass::make(p.hi, p.lo, false)};
}
std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}
bool expand_token(const std::string& tok, std::vector<int>& out) {  // range.rs:201-238
  std::string lower;
  for (char c : tok) lower.push_back((char)std::tolower((unsigned char)c));
  if (lower == "any" || lower == "random" || lower == "all") {
    for (int i = 0; i < 1326; i++) out.push_back(i);
    return true;
  }
  if (tok.size() == 4) {
    std::vector<int> cs = parse_cards(tok);
    if (cs.size() == 2) {
      if (cs[0] == cs[1]) return false;
      out.push_back(combo_index(cs[0], cs[1]));
      return true;
    }
  }
  std::vector<Pattern> pats;
  size_t dash = tok.find('-');
  if (dash != std::string::npos) {
    Pattern pa, pb;
    if (!pattern_parse(trim(tok.substr(0, dash)), pa) || !pattern_parse(trim(tok.substr(dash + 1)), pb)) return false;
    if (pa.hi == pa.lo && pb.hi == pb.lo) {
      int lo = std::min(pa.hi, pb.hi), hi = std::max(pa.hi, pb.hi);
      for (int r = lo; r <= hi; r++) pats.push_back({r, r, -1});
    } else if (pa.hi == pb.hi && pa.suited == pb.suited && pa.hi != pa.lo && pb.hi != pb.lo) {
      int lo = std::min(pa.lo, pb.lo), hi = std::max(pa.lo, pb.lo);
      for (int l = lo; l <= hi; l++) pats.push_back({pa.hi, l, pa.suited});
    } else {
      return false;
    }
  } else if (!tok.empty() && tok.back() == '+') {
    Pattern p;
    if (!pattern_parse(trim(tok.substr(0, tok.size() - 1)), p)) return false;
    if (p.hi == p.lo) {
      for (int r = p.hi; r <= 12; r++) pats.push_back({r, r, -1});
    } else {
      for (int l = p.lo; l < p.hi; l++) pats.push_back({p.hi, l, p.suited});
    }
  } else {
    Pattern p;
    if (!pattern_parse(tok, p)) return false;
    pats.push_back(p);
  }
  for (const Pattern& p : pats)
    for (const HandClass& c : pattern_classes(p))
      for (int k : c.combos()) out.push_back(k);
  return true;
}
}  // namespace

bool parse_range(const std::string& s, std::vector<float>& w) {  // range.rs:240-264
  w.assign(1326, 0.0f);
  size_t pos = 0;
  while (pos <= s.size()) {
    size_t comma = s.find(',', pos);
    if (comma == std::string::npos) comma = s.size();
    std::string raw = trim(s.substr(pos, comma - pos));
    pos = comma + 1;
    if (raw.empty()) { if (comma == s.size()) break; continue; }
    std::string tok = raw;
    float weight = 1.0f;
    size_t colon = raw.find(':');
    if (colon != std::string::npos) {
      tok = trim(raw.substr(0, colon));
      std::string ws = trim(raw.substr(colon + 1));
      char* end = nullptr;
      weight = std::strtof(ws.c_str(), &end);
      if (ws.empty() || *end) return false;
    }
    std::vector<int> cs;
    if (!expand_token(tok, cs)) return false;
    for (int c : cs) w[c] = weight;
    if (comma == s.size()) break;
  }
  return true;
}

// =================================================================== menu
bool parse_bet_size(const std::string& s, BetSize& out) {  // menu.rs:26-51
  std::string t;
  for (char c : trim(s)) t.push_back((char)std::tolower((unsigned char)c));
  if (t == "a" || t == "allin" || t == "all-in" || t == "all in" || t == "shove") { out = {BetSize::AllIn, 0}; return true; }
  if (!t.empty() && t.back() == '%') {
    std::string n = trim(t.substr(0, t.size() - 1));
    char* e = nullptr;
    double v = std::strtod(n.c_str(), &e);
    if (n.empty() || *e || v <= 0.0) return false;
    out = {BetSize::Pot, v / 100.0};
    return true;
  }
  if (!t.empty() && t.back() == 'x') {
    std::string n = trim(t.substr(0, t.size() - 1));
    char* e = nullptr;
    double v = std::strtod(n.c_str(), &e);
    if (n.empty() || *e || v <= 1.0) return false;
    out = {BetSize::Mult, v};
    return true;
  }
  return false;
}

PostflopMenu PostflopMenu::defaults() {  // menu.rs:121-133
  PostflopMenu m;
  m.flop.bet = {{BetSize::Pot, 0.33}, {BetSize::Pot, 0.75}};
  m.flop.raise = {{BetSize::Mult, 3.0}};
  m.turn.bet = {{BetSize::Pot, 0.5}, {BetSize::Pot, 1.0}};
  m.turn.raise = {{BetSize::Mult, 3.0}};
  m.river = m.turn;
  m.max_raises = 3;
  m.add_allin_threshold = 1.5;
  m.force_allin_threshold = 0.15;
  m.merge_threshold = 0.1;
  return m;
}

static inline Chips round_chips(double x) { return (Chips)std::round(x); }  // f64::round: half away from zero

static std::vector<std::pair<Chips, size_t>> finalize_tagged(const Situation& sit, std::vector<std::pair<Chips, size_t>> amounts,
                                                             double force_allin, double merge) {  // menu.rs:355-392
  Chips lo = std::min(sit.min_raise_to, sit.max_to);
  Chips pot_after_call = sit.pot + sit.to_call;
  std::vector<std::pair<Chips, size_t>> v;
  for (auto& at : amounts) {
    Chips a = std::min(std::max(at.first, lo), sit.max_to);
    Chips new_pot = pot_after_call + 2 * (a - sit.current_bet);
    Chips threshold = round_chips((double)new_pot * force_allin);
    std::pair<Chips, size_t> r = (sit.max_to - a <= threshold) ? std::make_pair(sit.max_to, at.second) : std::make_pair(a, at.second);
    if (r.first > sit.current_bet) v.push_back(r);
  }
  std::sort(v.begin(), v.end());
  std::vector<std::pair<Chips, size_t>> d;
  for (auto& x : v) if (d.empty() || d.back().first != x.first) d.push_back(x);
  v = d;
  if (merge > 0.0 && v.size() > 1) {
    auto ratio = [&](Chips a) { return (double)(a - sit.current_bet) / (double)pot_after_call; };
    std::vector<std::pair<Chips, size_t>> kept;
    for (auto it = v.rbegin(); it != v.rend(); ++it) {
      if (!kept.empty() && (1.0 + ratio(kept.back().first)) / (1.0 + ratio(it->first)) < 1.0 + merge) continue;
      kept.push_back(*it);
    }
    std::reverse(kept.begin(), kept.end());
    v = kept;
  }
  return v;
}

std::vector<Action> postflop_actions(const Situation& sit, const PostflopMenu& menu) {  // menu.rs:264-308
  std::vector<Action> out;
  if (sit.to_call > 0) {
    out.push_back({Fold, 0});
    out.push_back({Call, 0});
  } else {
    out.push_back({Check, 0});
  }
  if (!sit.can_raise() || sit.num_raises >= menu.max_raises) return out;
  const StreetSizes& sizes = menu.street(sit.street);
  const std::vector<BetSize>& list = sit.current_bet == 0 ? sizes.bet : sizes.raise;
  double pot_after_call = (double)(sit.pot + sit.to_call);
  std::vector<std::pair<Chips, size_t>> amounts;
  for (size_t i = 0; i < list.size(); i++) {
    const BetSize& b = list[i];
    Chips to;
    if (b.kind == BetSize::Pot) {
      to = sit.current_bet + round_chips(b.v * pot_after_call);
    } else if (b.kind == BetSize::Mult) {
      if (sit.current_bet == 0) continue;
      to = round_chips(b.v * (double)sit.current_bet);
    } else {
      to = sit.max_to;
    }
    amounts.push_back({to, i});
  }
  if (menu.add_allin_threshold > 0.0 && (double)(sit.max_to - sit.current_bet) <= menu.add_allin_threshold * pot_after_call)
    amounts.push_back({sit.max_to, list.size()});
  for (auto& a : finalize_tagged(sit, amounts, menu.force_allin_threshold, menu.merge_threshold)) out.push_back({Raise, a.first});
  return out;
}

// =================================================================== betting state
BettingState BettingState::street_start(Street st, Chips pot, Chips s0, Chips s1, Chips bb) {  // tree.rs:48-61
  BettingState b;
  b.street = st; b.pot = pot; b.commit[0] = b.commit[1] = 0; b.stack[0] = s0; b.stack[1] = s1;
  b.current_bet = 0; b.last_raise_inc = bb; b.num_raises = 0; b.to_act = 0;
  b.needs[0] = s0 > 0; b.needs[1] = s1 > 0;
  return b;
}

Situation BettingState::situation(Chips bb) const {  // tree.rs:63-77
  int p = to_act, o = p ^ 1;
  Situation s;
  s.street = street; s.pot = pot;
  s.to_call = std::min(std::max(current_bet - commit[p], (Chips)0), stack[p]);
  s.current_bet = current_bet;
  s.max_to = std::min(commit[p] + stack[p], commit[o] + stack[o]);
  s.min_raise_to = current_bet + std::max(last_raise_inc, bb);
  s.num_raises = num_raises; s.limpers = 0; s.big_blind = bb;
  return s;
}

Transition apply(const BettingState& st, const Action& a) {  // tree.rs:93-125
  int p = st.to_act, o = p ^ 1;
  BettingState s = st;
  Chips put = 0;
  if (a.kind == Fold) return {FoldT, p, st, 0};
  if (a.kind == Call) {
    put = std::min(std::max(s.current_bet - s.commit[p], (Chips)0), s.stack[p]);
  } else if (a.kind == Raise) {
    put = a.to - s.commit[p];
    Chips inc = a.to - s.current_bet;
    if (inc >= s.last_raise_inc) s.last_raise_inc = inc;
    s.current_bet = a.to;
    s.num_raises += 1;
    s.needs[o] = s.stack[o] > 0;
  }
  s.commit[p] += put;
  s.stack[p] -= put;
  s.pot += put;
  s.needs[p] = false;
  if (s.needs[o]) {
    s.to_act = o;
    return {Continue, -1, s, put};
  }
  return {StreetOver, -1, s, put};
}

// =================================================================== iso
const int PAIRS[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};  // iso.rs:174
static int pair_index(int a, int b) {
  if (a > b) std::swap(a, b);
  for (int i = 0; i < 6; i++) if (PAIRS[i][0] == a && PAIRS[i][1] == b) return i;
  return -1;
}
static int swap_card(int c, int a, int b) {  // iso.rs:182-192
  int s = suit_of(c);
  int s2 = s == a ? b : s == b ? a : s;
  return rank_of(c) * 4 + s2;
}
static void suit_ranks(const std::vector<int>& board, uint16_t m[4]) {
  m[0] = m[1] = m[2] = m[3] = 0;
  for (int c : board) m[suit_of(c)] |= (uint16_t)(1u << rank_of(c));
}

Isomorphism make_iso(const HandSet& hs, bool enabled) {  // iso.rs:245-273
  Isomorphism iso;
  iso.enabled = enabled;
  if (!enabled) return iso;
  for (int pi = 0; pi < 6; pi++) {
    int a = PAIRS[pi][0], b = PAIRS[pi][1];
    std::vector<uint16_t> pp[2];
    bool ok = true;
    for (int p = 0; p < 2 && ok; p++) {
      const std::vector<float>& w = hs.weights[p];
      float max_w = 0.0f;
      for (float x : w) max_w = std::max(max_w, x);
      max_w = std::max(max_w, 1e-30f);
      for (size_t h = 0; h < hs.len(p); h++) {
        int s1 = swap_card(hs.cards[p][h].first, a, b), s2 = swap_card(hs.cards[p][h].second, a, b);
        uint32_t j = hs.lookup[p][combo_index(s1, s2)];
        if (j == NONE) { ok = false; break; }
        if (std::fabs(w[h] - w[j]) > 1e-4f * max_w) { ok = false; break; }
        pp[p].push_back((uint16_t)j);
      }
    }
    if (!ok) continue;
    iso.range_symmetric[pi] = true;
    iso.perm[0][pi] = pp[0];
    iso.perm[1][pi] = pp[1];
  }
  return iso;
}

void Isomorphism::chance_cards(const std::vector<int>& board, const std::vector<int>& available, std::vector<int>& canon,
                               std::vector<std::array<int, 3>>& refs) const {  // iso.rs:277-300
  canon.clear();
  refs.clear();
  bool any = false;
  for (bool b : range_symmetric) any |= b;
  if (!enabled || !any) { canon = available; return; }
  uint16_t m[4];
  suit_ranks(board, m);
  for (int c : available) {
    int s = suit_of(c), found = -1;
    for (int s0 = 0; s0 < s; s0++) {
      int pi = pair_index(s0, s);
      if (range_symmetric[pi] && m[s0] == m[s]) { found = s0; break; }
    }
    if (found >= 0) refs.push_back({c, rank_of(c) * 4 + found, pair_index(found, s)});
    else canon.push_back(c);
  }
}

// =================================================================== tree
std::vector<Action> Tree::actions(size_t id) const {
  std::vector<Action> out;
  const Node& n = nodes[id];
  for (uint32_t c = n.first_child; c < n.first_child + n.num_children; c++) out.push_back(nodes[c].edge_action);
  return out;
}

// ---- rand 0.9.5's StdRng and the sampled chance cards. The crate is not in
// the reference tree (Cargo.lock pins rand 0.9.5, rand_chacha 0.9.0,
// rand_core 0.9.5); restated from the published algorithms:
//   StdRng = ChaCha12Rng: the ChaCha block function (D. J. Bernstein; RFC
//     8439's constants and key layout) with 12 rounds, a 64-bit block
//     counter in words 12-13 and stream 0 in words 14-15; 16-word blocks in
//     counter order, read as u32 words; next_u64 = two words, the first as
//     the low half (rand_core BlockRng);
//   seed_from_u64: PCG32 (M. O'Neill, XSH RR output; multiplier
//     6364136223846793005, increment 11634580027462260723), the state
//     advanced before each of the eight 32-bit key words;
//   random::<f64>() = (next_u64 >> 11) * 2^-53.
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
void chacha_block(int rounds, const uint32_t in[16], uint32_t out[16]) {
  uint32_t x[16];
  std::memcpy(x, in, sizeof x);
  auto qr = [&](int a, int b, int c, int d) {
    x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 16);
    x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 12);
    x[a] += x[b]; x[d] = rotl32(x[d] ^ x[a], 8);
    x[c] += x[d]; x[b] = rotl32(x[b] ^ x[c], 7);
  };
  for (int i = 0; i < rounds; i += 2) {
    qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15);
    qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14);
  }
  for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

uint32_t pcg32_next(uint64_t& state, uint64_t inc) {
  state = state * 6364136223846793005ull + inc;
  const uint32_t xs = (uint32_t)(((state >> 18) ^ state) >> 27);
  const uint32_t rot = (uint32_t)(state >> 59);
  return (xs >> rot) | (xs << ((32 - rot) & 31));
}
StdRng::StdRng(uint64_t state) {
  const uint64_t INC = 11634580027462260723ull;
  for (int i = 0; i < 8; i++) key[i] = pcg32_next(state, INC);
}
StdRng::StdRng(const uint32_t k[8], int r) : rounds(r) { std::memcpy(key, k, sizeof key); }
void StdRng::refill() {
  for (int b = 0; b < 4; b++) {
    const uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                             key[4], key[5], key[6], key[7], (uint32_t)counter, (uint32_t)(counter >> 32), 0u, 0u};
    chacha_block(rounds, in, buf + 16 * b);
    counter++;
  }
}
uint32_t StdRng::next_u32() {
  if (idx >= 64) {
    refill();
    idx = 0;
  }
  return buf[idx++];
}
uint64_t StdRng::next_u64() {
  if (idx < 63) {
    const uint64_t v = (uint64_t)buf[idx + 1] << 32 | buf[idx];
    idx += 2;
    return v;
  }
  if (idx >= 64) {
    refill();
    idx = 2;
    return (uint64_t)buf[1] << 32 | buf[0];
  }
  const uint64_t x = buf[63];
  refill();
  idx = 1;
  return (uint64_t)buf[0] << 32 | x;
}
double StdRng::next_f64() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

// sampling.rs:9-56: k of the available cards, every suit in proportion to its
// remaining cards (largest remainder, ties broken by a random draw), ranks
// spread evenly within a suit from a random offset; sorted.
std::vector<int> sample_cards(const std::vector<int>& available, size_t k, StdRng& rng) {
  if (k >= available.size()) return available;
  std::vector<int> by_suit[4];
  for (int c : available) by_suit[c & 3].push_back(c);
  for (auto& v : by_suit) std::sort(v.begin(), v.end(), [](int a, int b) { return (a >> 2) < (b >> 2); });
  const double total = (double)available.size();
  size_t quotas[4];
  std::vector<std::tuple<double, double, int>> fracs;
  for (int su = 0; su < 4; su++) {
    const double exact = (double)k * (double)by_suit[su].size() / total;
    quotas[su] = (size_t)std::floor(exact);
    fracs.emplace_back(exact - std::floor(exact), rng.next_f64(), su);
  }
  size_t assigned = quotas[0] + quotas[1] + quotas[2] + quotas[3];
  std::sort(fracs.begin(), fracs.end(), [](const auto& a, const auto& b) { return b < a; });
  for (const auto& f : fracs) {
    if (assigned >= k) break;
    const int su = std::get<2>(f);
    if (quotas[su] < by_suit[su].size()) {
      quotas[su]++;
      assigned++;
    }
  }
  std::vector<int> out;
  for (int su = 0; su < 4; su++) {
    const size_t m = by_suit[su].size(), q = quotas[su];
    if (q == 0) continue;
    const double step = (double)m / (double)q;
    const double offset = rng.next_f64() * step;
    for (size_t j = 0; j < q; j++) {
      const size_t idx = std::min((size_t)std::floor(offset + (double)j * step), m - 1);
      out.push_back(by_suit[su][idx]);
    }
  }
  std::sort(out.begin(), out.end());
  out.erase(std::unique(out.begin(), out.end()), out.end());
  return out;
}

namespace {
struct TreeParams {  // tree.rs:276-335
  Street root_street;
  const std::vector<int>* root_board;
  const PostflopMenu* menu;
  const PostflopMenu* future_menu;
  Chips big_blind;
  size_t num_hands[2];
  double rake_rate, rake_cap;
  const Isomorphism* iso;
  int chance_mode = 0;  // tree.rs:128-136
  size_t sample_turn = 0, sample_river = 0;
  const PostflopMenu& menu_for(Street s) const { return s == root_street ? *menu : *future_menu; }
  std::vector<Action> actions(const BettingState& st) const { return postflop_actions(st.situation(big_blind), menu_for(st.street)); }
  float rake(Chips pot) const { return (float)std::min((double)pot * rake_rate, rake_cap); }
  // tree.rs:314-334: a sample of the available cards when sampling applies
  // (not an all-in runout, fewer cards than available), else every card up
  // to suit isomorphism
  void chance_cards(const std::vector<int>& dealt, bool all_in, Street next, StdRng& rng, std::vector<int>& canon,
                    std::vector<std::array<int, 3>>& refs, size_t& n_avail) const {
    std::vector<int> board = *root_board;
    board.insert(board.end(), dealt.begin(), dealt.end());
    uint64_t used = 0;
    for (int c : board) used |= card_bit(c);
    std::vector<int> available;
    for (int c = 0; c < 52; c++) if (!(used & card_bit(c))) available.push_back(c);
    n_avail = available.size();
    const size_t k = next == Turn ? sample_turn : sample_river;
    if (chance_mode == 1 && !all_in && k < n_avail) {
      canon = sample_cards(available, k, rng);
      refs.clear();
      return;
    }
    iso->chance_cards(board, available, canon, refs);
  }
};

void count_after_street(const TreeParams& P, const BettingState& st, std::vector<int>& dealt, StdRng& rng, TreeStats& stats);
void count_player(const TreeParams& P, const BettingState& st, std::vector<int>& dealt, StdRng& rng, TreeStats& stats) {  // tree.rs:345-362
  stats.player_nodes += 1;
  std::vector<Action> acts = P.actions(st);
  stats.storage_cells += (uint64_t)(acts.size() * P.num_hands[st.to_act]);
  for (const Action& a : acts) {
    Transition t = apply(st, a);
    if (t.kind == Continue) count_player(P, t.st, dealt, rng, stats);
    else if (t.kind == FoldT) stats.terminal_nodes += 1;
    else count_after_street(P, t.st, dealt, rng, stats);
  }
}
void count_after_street(const TreeParams& P, const BettingState& st, std::vector<int>& dealt, StdRng& rng, TreeStats& stats) {  // tree.rs:364-401
  if (st.street == River) { stats.terminal_nodes += 1; return; }
  stats.chance_nodes += 1;
  bool all_in = st.stack[0] == 0 || st.stack[1] == 0;
  Street next = (Street)(st.street + 1);
  std::vector<int> cards;
  std::vector<std::array<int, 3>> refs;
  size_t n_avail;
  P.chance_cards(dealt, all_in, next, rng, cards, refs, n_avail);
  bool uniform = refs.empty() && !P.iso->enabled;
  size_t reps = uniform ? 1 : cards.size();
  TreeStats sub;
  for (size_t i = 0; i < reps; i++) {
    dealt.push_back(cards[i]);
    if (all_in) {
      BettingState s2 = st;
      s2.street = next;
      count_after_street(P, s2, dealt, rng, sub);
    } else {
      BettingState s2 = BettingState::street_start(next, st.pot, st.stack[0], st.stack[1], P.big_blind);
      count_player(P, s2, dealt, rng, sub);
    }
    dealt.pop_back();
  }
  uint64_t k = uniform ? cards.size() : 1;
  stats.player_nodes += sub.player_nodes * k;
  stats.chance_nodes += sub.chance_nodes * k;
  stats.terminal_nodes += sub.terminal_nodes * k;
  stats.storage_cells += sub.storage_cells * k;
}

Node placeholder() {
  Node n;
  n.kind = KShowdown; n.street = River; n.num_children = 0; n.first_child = 0;
  n.invest[0] = n.invest[1] = 0; n.offset = 0; n.slot = 0; n.rake = 0.0f; n.chance_factor = 1.0f;
  n.edge_kind = 0; n.edge_action = {Check, 0}; n.edge_card = -1;
  return n;
}

struct Builder {  // tree.rs:403-574
  const TreeParams& P;
  Tree t;
  std::unordered_map<uint64_t, uint32_t> board_index;
  StdRng rng;  // tree.rs:415-429: seeded with the sampled mode's seed, else 0
  Builder(const TreeParams& p, uint64_t seed) : P(p), rng(seed) {}
  size_t alloc(size_t k) { size_t f = t.nodes.size(); t.nodes.resize(f + k, placeholder()); return f; }
  void fill_player(size_t id, const BettingState& st, const int32_t invest[2], std::vector<int>& dealt) {
    std::vector<Action> acts = P.actions(st);
    size_t k = acts.size();
    size_t first = alloc(k);
    Node& n = t.nodes[id];
    n.kind = st.to_act == 0 ? KPlayer0 : KPlayer1;
    n.street = st.street; n.num_children = (uint16_t)k; n.first_child = (uint32_t)first;
    n.invest[0] = invest[0]; n.invest[1] = invest[1];
    n.offset = t.storage_len; n.slot = t.num_player_nodes; n.rake = 0.0f; n.chance_factor = 1.0f;
    t.storage_len += (uint64_t)(k * P.num_hands[st.to_act]);
    t.num_player_nodes += 1;
    for (size_t i = 0; i < k; i++) {
      size_t child = first + i;
      Transition tr = apply(st, acts[i]);
      int32_t inv[2] = {invest[0], invest[1]};
      inv[st.to_act] += (int32_t)tr.put;
      if (tr.kind == Continue) {
        t.nodes[child].edge_kind = 1; t.nodes[child].edge_action = acts[i];
        fill_player(child, tr.st, inv, dealt);
      } else if (tr.kind == FoldT) {
        int f = tr.folder, o = f ^ 1;
        Chips uncalled = std::max(st.commit[o] - st.commit[f], (Chips)0);
        Node c = placeholder();
        c.kind = f == 0 ? KFold0 : KFold1;
        c.street = st.street;
        c.invest[0] = inv[0]; c.invest[1] = inv[1];
        c.rake = P.rake(st.pot - uncalled);
        c.edge_kind = 1; c.edge_action = acts[i];
        t.nodes[child] = c;
      } else {
        t.nodes[child].edge_kind = 1; t.nodes[child].edge_action = acts[i];
        fill_after_street(child, tr.st, inv, dealt);
      }
    }
  }
  void fill_after_street(size_t id, const BettingState& st, const int32_t invest[2], std::vector<int>& dealt) {
    uint8_t ek = t.nodes[id].edge_kind; Action ea = t.nodes[id].edge_action; int ec = t.nodes[id].edge_card;
    if (st.street == River) {
      uint32_t table = table_for(dealt);
      Node n = placeholder();
      n.kind = KShowdown; n.street = River; n.invest[0] = invest[0]; n.invest[1] = invest[1];
      n.offset = table; n.rake = P.rake(st.pot);
      n.edge_kind = ek; n.edge_action = ea; n.edge_card = ec;
      t.nodes[id] = n;
      return;
    }
    bool all_in = st.stack[0] == 0 || st.stack[1] == 0;
    Street next = (Street)(st.street + 1);
    std::vector<int> cards;
    std::vector<std::array<int, 3>> refs;
    size_t n_avail;
    P.chance_cards(dealt, all_in, next, rng, cards, refs, n_avail);
    size_t blen = P.root_board->size() + dealt.size();
    float compatible = (float)(52 - blen - 4);
    size_t total_cards = cards.size() + refs.size();
    size_t first = alloc(cards.size());
    uint32_t slot = (uint32_t)t.chance.size();
    ChanceInfo info;
    for (auto& r : refs) {
      size_t k = std::find(cards.begin(), cards.end(), r[1]) - cards.begin();
      info.iso.push_back({(uint16_t)k, (uint8_t)r[2]});
      info.iso_cards.push_back(r[0]);
    }
    t.chance.push_back(info);
    Node n = placeholder();
    n.kind = KChance; n.street = st.street; n.num_children = (uint16_t)cards.size(); n.first_child = (uint32_t)first;
    n.invest[0] = invest[0]; n.invest[1] = invest[1]; n.slot = slot;
    n.chance_factor = 1.0f / ((float)total_cards * compatible / (float)n_avail);
    n.edge_kind = ek; n.edge_action = ea; n.edge_card = ec;
    t.nodes[id] = n;
    for (size_t i = 0; i < cards.size(); i++) {
      size_t child = first + i;
      t.nodes[child].edge_kind = 2; t.nodes[child].edge_card = cards[i];
      dealt.push_back(cards[i]);
      if (all_in) {
        BettingState s2 = st;
        s2.street = next;
        fill_after_street(child, s2, invest, dealt);
      } else {
        BettingState s2 = BettingState::street_start(next, st.pot, st.stack[0], st.stack[1], P.big_blind);
        fill_player(child, s2, invest, dealt);
      }
      dealt.pop_back();
    }
  }
  uint32_t table_for(const std::vector<int>& dealt) {
    uint64_t key = 0;
    for (int c : dealt) key |= card_bit(c);
    auto it = board_index.find(key);
    if (it != board_index.end()) return it->second;
    uint32_t i = (uint32_t)t.boards.size();
    std::vector<int> b = *P.root_board;
    b.insert(b.end(), dealt.begin(), dealt.end());
    t.boards.push_back(b);
    board_index[key] = i;
    return i;
  }
};
}  // namespace

// =================================================================== hands
HandSet make_handset(const std::vector<float>* ranges, uint64_t board, float prune, const std::vector<int>* keep) {  // hands.rs:35-76
  HandSet hs;
  for (int p = 0; p < 2; p++) {
    float max_w = 0.0f;
    for (float x : ranges[p]) max_w = std::max(max_w, x);
    float cutoff = max_w * prune;
    for (int c = 0; c < 1326; c++) {
      const ComboCards& cc = combo_cards(c);
      if ((board & card_bit(cc.hi)) || (board & card_bit(cc.lo))) continue;
      float w = ranges[p][c];
      bool forced = keep && std::find(keep[p].begin(), keep[p].end(), c) != keep[p].end();
      if ((w > 0.0f && w >= cutoff) || (forced && w > 0.0f)) {
        hs.combos[p].push_back((uint16_t)c);
        hs.cards[p].push_back({cc.hi, cc.lo});
        hs.weights[p].push_back(w);
      }
    }
  }
  for (int p = 0; p < 2; p++) {
    hs.lookup[p].assign(1326, NONE);
    for (size_t i = 0; i < hs.combos[p].size(); i++) hs.lookup[p][hs.combos[p][i]] = (uint32_t)i;
  }
  for (int p = 0; p < 2; p++) {
    for (uint16_t c : hs.combos[p]) hs.same[p].push_back(hs.lookup[p ^ 1][c]);
    hs.blocked[p].assign(52, {});
    for (size_t i = 0; i < hs.cards[p].size(); i++) {
      hs.blocked[p][hs.cards[p][i].first].push_back((uint32_t)i);
      hs.blocked[p][hs.cards[p][i].second].push_back((uint32_t)i);
    }
  }
  // hands.rs:95-98: sum_i compat_0(i) * w0(i)
  std::vector<double> compat = hs.compatible_reach(0, hs.weights[1].data());
  double s = 0.0;
  for (size_t i = 0; i < compat.size(); i++) s += compat[i] * (double)hs.weights[0][i];
  hs.num_combinations = s;
  return hs;
}

std::vector<double> HandSet::compatible_reach(int p, const float* opp) const {  // hands.rs:102-125
  int o = p ^ 1;
  double total = 0.0;
  double minus[52] = {0};
  for (size_t k = 0; k < len(o); k++) {
    float w = opp[k];
    if (w != 0.0f) {
      double wd = (double)w;
      total += wd;
      minus[cards[o][k].first] += wd;
      minus[cards[o][k].second] += wd;
    }
  }
  std::vector<double> out(len(p));
  for (size_t i = 0; i < len(p); i++) {
    double s = same[p][i] == NONE ? 0.0 : (double)opp[same[p][i]];
    out[i] = std::max(total - minus[cards[p][i].first] - minus[cards[p][i].second] + s, 0.0);
  }
  return out;
}

StrengthTable make_table(const HandSet& hs, const std::vector<int>& board) {  // hands.rs:137-158
  StrengthTable t;
  uint64_t mask = 0, bset = 0;
  for (int c : board) { mask |= eval_mask_bit(c); bset |= card_bit(c); }
  for (int p = 0; p < 2; p++) {
    for (size_t i = 0; i < hs.len(p); i++) {
      int a = hs.cards[p][i].first, b = hs.cards[p][i].second;
      if ((bset & card_bit(a)) || (bset & card_bit(b))) continue;
      t.lists[p].push_back({evaluate_mask(mask | eval_mask_bit(a) | eval_mask_bit(b)), (uint16_t)i, (uint8_t)a, (uint8_t)b});
    }
    std::sort(t.lists[p].begin(), t.lists[p].end());
  }
  return t;
}

// =================================================================== dcfr
Discount make_discount(uint32_t t, const DcfrParams& p) {  // cfr.rs:86-106
  uint32_t tm = t < 1 ? 1 : t;
  double prev = (double)(tm - 1);
  double pa = std::pow(prev, p.alpha);
  double tf = (double)tm;
  float neg = 0.0f;
  if (p.beta_some) {
    double pb = std::pow(prev, p.beta);
    neg = (float)(pb / (pb + 1.0));
  }
  double since = tf - 1.0;
  if (p.restart_average) {
    uint32_t t0 = tm - 1;
    uint32_t p4 = t0 == 0 ? 0 : (1u << ((31 - __builtin_clz(t0)) & ~1u));
    since = (double)(t0 - p4);
  }
  float strat = (float)std::pow(since / (since + 1.0), p.gamma);
  return {(float)(pa / (pa + 1.0)), neg, strat, !p.unweighted_average};
}

template <typename V> static inline V recip(V d) {  // cfr.rs:113-121
  V i = V(1) / d;
  return (d > V(0) && std::isfinite(i)) ? i : V(0);
}
template <typename V> static inline V pos0(V x) { return x > V(0) ? x : V(0); }  // max(x, 0)

// cfr.rs:124-154 (sums loaded from f32 storage into V)
template <typename V> static void normalize_in_place(V* s, size_t na, size_t n) {
  if (n == 0) return;
  std::vector<V> inv(s, s + n);
  for (size_t a = 1; a < na; a++)
    for (size_t h = 0; h < n; h++) inv[h] += s[a * n + h];
  bool any_zero = false;
  for (size_t h = 0; h < n; h++) { inv[h] = recip(inv[h]); any_zero |= inv[h] == V(0); }
  for (size_t a = 0; a < na; a++)
    for (size_t h = 0; h < n; h++) s[a * n + h] *= inv[h];
  if (any_zero) {
    V uniform = V(1) / (V)na;
    for (size_t h = 0; h < n; h++)
      if (inv[h] == V(0)) for (size_t a = 0; a < na; a++) s[a * n + h] = uniform;
  }
}

// cfr.rs:156-207
static void regret_match_into(const float* r, size_t na, size_t n, float* out) {
  if (n == 0) return;
  std::vector<float> inv(n);
  for (size_t h = 0; h < n; h++) inv[h] = pos0(r[h]);
  for (size_t a = 1; a < na; a++)
    for (size_t h = 0; h < n; h++) inv[h] += pos0(r[a * n + h]);
  bool any_zero = false;
  for (size_t h = 0; h < n; h++) { inv[h] = recip(inv[h]); any_zero |= inv[h] == 0.0f; }
  for (size_t a = 0; a < na; a++)
    for (size_t h = 0; h < n; h++) out[a * n + h] = pos0(r[a * n + h]) * inv[h];
  if (any_zero) {
    float uniform = 1.0f / (float)na;
    for (size_t h = 0; h < n; h++)
      if (inv[h] == 0.0f) for (size_t a = 0; a < na; a++) out[a * n + h] = uniform;
  }
}

static double fold_payoff(const Ctx& ctx, const Node& node, int p) {  // cfr.rs:224-235
  int f = node.kind == KFold0 ? 0 : 1;
  double v = f == p ? -(ctx.half_pot + (double)node.invest[p]) : ctx.half_pot + (double)node.invest[f] - (double)node.rake;
  return v * ctx.norm;
}
static void showdown_payoffs(const Ctx& ctx, const Node& node, int p, double& win, double& lose, double& tie) {  // cfr.rs:238-246
  int o = p ^ 1;
  double rake = (double)node.rake;
  win = (ctx.half_pot + (double)node.invest[o] - rake) * ctx.norm;
  lose = -(ctx.half_pot + (double)node.invest[p]) * ctx.norm;
  tie = 0.5 * ((double)(node.invest[o] - node.invest[p]) - rake) * ctx.norm;
}

// =================================================================== terminal evaluators (eval.rs)
template <typename V> static void eval_fold_t(const HandSet& hs, int p, const V* opp, double payoff, V* out) {  // eval.rs:40-64
  int o = p ^ 1;
  double total = 0.0, minus[52] = {0};
  for (size_t k = 0; k < hs.len(o); k++) {
    V w = opp[k];
    if (w != V(0)) {
      double wd = (double)w;
      total += wd;
      minus[hs.cards[o][k].first] += wd;
      minus[hs.cards[o][k].second] += wd;
    }
  }
  size_t n = hs.len(p);
  if (total == 0.0) { for (size_t i = 0; i < n; i++) out[i] = V(0); return; }
  for (size_t i = 0; i < n; i++) {
    uint32_t sm = hs.same[p][i];
    double s = sm == NONE ? 0.0 : (double)opp[sm];
    double m2 = minus[hs.cards[p][i].first] + minus[hs.cards[p][i].second];
    out[i] = (V)(payoff * (total - m2 + s));
  }
}

template <typename V> static void add_ties_t(const HandSet& hs, const StrengthTable& t, int p, const V* opp, double tie, V* out) {  // eval.rs:347-379
  int o = p ^ 1;
  const std::vector<Entry>& mine = t.lists[p];
  const std::vector<Entry>& theirs = t.lists[o];
  size_t j = 0, k = 0;
  while (k < mine.size()) {
    uint32_t s = mine[k].strength;
    while (j < theirs.size() && theirs[j].strength < s) j++;
    double eq_sum = 0.0, eq_minus[52] = {0};
    size_t jj = j;
    while (jj < theirs.size() && theirs[jj].strength == s) {
      const Entry& e = theirs[jj];
      V w = opp[e.idx];
      if (w != V(0)) { double wd = (double)w; eq_sum += wd; eq_minus[e.c1] += wd; eq_minus[e.c2] += wd; }
      jj++;
    }
    while (k < mine.size() && mine[k].strength == s) {
      const Entry& e = mine[k];
      uint32_t q = hs.same[p][e.idx];
      double same_w = q == NONE ? 0.0 : (double)opp[q];
      double eq = eq_sum - (eq_minus[e.c1] + eq_minus[e.c2]) + same_w;
      out[e.idx] += (V)(tie * eq);
      k++;
    }
  }
}

template <typename V> static void eval_showdown_t(const HandSet& hs, const StrengthTable& t, int p, const V* opp,
                                                  double win, double lose, double tie, V* out) {  // eval.rs:69-148
  size_t n = hs.len(p);
  for (size_t i = 0; i < n; i++) out[i] = V(0);
  int o = p ^ 1;
  const std::vector<Entry>& mine = t.lists[p];
  const std::vector<Entry>& theirs = t.lists[o];
  double sum = 0.0, minus[52] = {0};
  size_t j = 0, k = 0;
  while (k < mine.size()) {
    uint32_t s = mine[k].strength;
    while (j < theirs.size()) {
      const Entry& e = theirs[j];
      if (e.strength >= s) break;
      V w = opp[e.idx];
      if (w != V(0)) { double wd = (double)w; sum += wd; minus[e.c1] += wd; minus[e.c2] += wd; }
      j++;
    }
    while (k < mine.size()) {
      const Entry& e = mine[k];
      if (e.strength != s) break;
      out[e.idx] = (V)(win * (sum - (minus[e.c1] + minus[e.c2])));
      k++;
    }
  }
  sum = 0.0;
  std::memset(minus, 0, sizeof(minus));
  j = theirs.size();
  k = mine.size();
  while (k > 0) {
    uint32_t s = mine[k - 1].strength;
    while (j > 0) {
      const Entry& e = theirs[j - 1];
      if (e.strength <= s) break;
      V w = opp[e.idx];
      if (w != V(0)) { double wd = (double)w; sum += wd; minus[e.c1] += wd; minus[e.c2] += wd; }
      j--;
    }
    while (k > 0) {
      const Entry& e = mine[k - 1];
      if (e.strength != s) break;
      out[e.idx] += (V)(lose * (sum - (minus[e.c1] + minus[e.c2])));
      k--;
    }
  }
  if (tie != 0.0) add_ties_t(hs, t, p, opp, tie, out);
}

constexpr int LANES = 4;  // eval.rs:150-151

template <typename V> static void fold_batch_t(const HandSet& hs, int p, int k, const V* const* reaches, const double* payoffs,
                                               V* const* outs) {  // eval.rs:172-206
  int o = p ^ 1;
  double total[LANES] = {0}, minus[52][LANES];
  std::memset(minus, 0, sizeof(minus));
  for (size_t j = 0; j < hs.len(o); j++) {
    double w[LANES] = {0};
    bool nz = false;
    for (int t = 0; t < k; t++) { w[t] = (double)reaches[t][j]; nz |= w[t] != 0.0; }
    if (!nz) continue;
    int a = hs.cards[o][j].first, b = hs.cards[o][j].second;
    for (int t = 0; t < LANES; t++) { total[t] += w[t]; minus[a][t] += w[t]; minus[b][t] += w[t]; }
  }
  for (size_t h = 0; h < hs.len(p); h++) {
    int a = hs.cards[p][h].first, b = hs.cards[p][h].second;
    uint32_t s = hs.same[p][h];
    for (int t = 0; t < k; t++) {
      double sw = s == NONE ? 0.0 : (double)reaches[t][s];
      outs[t][h] = (V)(payoffs[t] * (total[t] - minus[a][t] - minus[b][t] + sw));
    }
  }
}

template <typename V> static void showdown_batch_t(const HandSet& hs, const StrengthTable& tb, int p, int k, const V* const* reaches,
                                                   const double* win, const double* lose, const double* tie, V* const* outs,
                                                   const int* has_fold, const double* fold_pay, V* const* fold_outs) {  // eval.rs:214-344
  int o = p ^ 1;
  const std::vector<Entry>& mine = tb.lists[p];
  const std::vector<Entry>& theirs = tb.lists[o];
  double wv[LANES] = {0}, lv[LANES] = {0};
  for (int t = 0; t < k; t++) { wv[t] = win[t]; lv[t] = lose[t]; }
  std::vector<std::array<double, LANES>> inter(theirs.size()), res(mine.size());
  for (size_t j = 0; j < theirs.size(); j++) {
    std::array<double, LANES> w{};
    for (int t = 0; t < k; t++) w[t] = (double)reaches[t][theirs[j].idx];
    inter[j] = w;
  }
  for (auto& r : res) r.fill(0.0);
  auto nonzero = [](const std::array<double, LANES>& w) { for (double x : w) if (x != 0.0) return true; return false; };
  double sum[LANES] = {0}, minus[52][LANES];
  std::memset(minus, 0, sizeof(minus));
  size_t i = 0, j = 0;
  while (i < mine.size()) {
    uint32_t s = mine[i].strength;
    while (j < theirs.size() && theirs[j].strength < s) {
      if (nonzero(inter[j])) {
        const Entry& e = theirs[j];
        for (int t = 0; t < LANES; t++) { sum[t] += inter[j][t]; minus[e.c1][t] += inter[j][t]; minus[e.c2][t] += inter[j][t]; }
      }
      j++;
    }
    while (i < mine.size() && mine[i].strength == s) {
      const Entry& e = mine[i];
      for (int t = 0; t < LANES; t++) res[i][t] = wv[t] * (sum[t] - minus[e.c1][t] - minus[e.c2][t]);
      i++;
    }
  }
  bool any_fold = false;
  for (int t = 0; t < k; t++) any_fold |= has_fold[t] != 0;
  if (any_fold) {
    while (j < theirs.size()) {
      if (nonzero(inter[j])) {
        const Entry& e = theirs[j];
        for (int t = 0; t < LANES; t++) { sum[t] += inter[j][t]; minus[e.c1][t] += inter[j][t]; minus[e.c2][t] += inter[j][t]; }
      }
      j++;
    }
    for (int t = 0; t < k; t++) {
      if (!has_fold[t]) continue;
      for (size_t h = 0; h < hs.len(p); h++) {
        int a = hs.cards[p][h].first, b = hs.cards[p][h].second;
        uint32_t sm = hs.same[p][h];
        double sw = sm == NONE ? 0.0 : (double)reaches[t][sm];
        fold_outs[t][h] = (V)(fold_pay[t] * (sum[t] - minus[a][t] - minus[b][t] + sw));
      }
    }
  }
  double sum2[LANES] = {0}, minus2[52][LANES];
  std::memset(minus2, 0, sizeof(minus2));
  i = mine.size();
  j = theirs.size();
  while (i > 0) {
    uint32_t s = mine[i - 1].strength;
    while (j > 0 && theirs[j - 1].strength > s) {
      if (nonzero(inter[j - 1])) {
        const Entry& e = theirs[j - 1];
        for (int t = 0; t < LANES; t++) { sum2[t] += inter[j - 1][t]; minus2[e.c1][t] += inter[j - 1][t]; minus2[e.c2][t] += inter[j - 1][t]; }
      }
      j--;
    }
    while (i > 0 && mine[i - 1].strength == s) {
      const Entry& e = mine[i - 1];
      for (int t = 0; t < LANES; t++) res[i - 1][t] += lv[t] * (sum2[t] - minus2[e.c1][t] - minus2[e.c2][t]);
      i--;
    }
  }
  for (size_t m = 0; m < mine.size(); m++)
    for (int t = 0; t < k; t++) outs[t][mine[m].idx] = (V)res[m][t];
  for (int t = 0; t < k; t++)
    if (tie[t] != 0.0) add_ties_t(hs, tb, p, reaches[t], tie[t], outs[t]);
}

void eval_fold(const HandSet& hs, int p, const float* r, double pay, float* out) { eval_fold_t<float>(hs, p, r, pay, out); }
void eval_showdown(const HandSet& hs, const StrengthTable& t, int p, const float* r, double w, double l, double ti, float* out) {
  eval_showdown_t<float>(hs, t, p, r, w, l, ti, out);
}
void fold_batch(const HandSet& hs, int p, int k, const float* const* r, const double* pay, float* const* outs) {
  fold_batch_t<float>(hs, p, k, r, pay, outs);
}
void showdown_batch(const HandSet& hs, const StrengthTable& t, int p, int k, const float* const* r, const double* w, const double* l,
                    const double* ti, float* const* outs, const int* hf, const double* fp, float* const* fo) {
  showdown_batch_t<float>(hs, t, p, k, r, w, l, ti, outs, hf, fp, fo);
}

// =================================================================== traversal (cfr.rs, river.rs)
template <typename V> static bool all_zero(const V* x, size_t n) {
  for (size_t i = 0; i < n; i++) if (x[i] != V(0)) return false;
  return true;
}

template <typename V> static void traverse_t(const Ctx& ctx, size_t id, int p, const V* opp, size_t nq, Mode mode, V* out, size_t n);

// ---- 16-bit storage (storage.rs:108-128, 236-330)
// Rust's saturating `f32 as i16` / `as u16`: truncation toward zero, NaN to 0.
static int16_t sat_i16(float y) {
  if (!(y == y)) return 0;
  if (y >= 32767.0f) return 32767;
  if (y <= -32768.0f) return -32768;
  return (int16_t)(int32_t)y;
}
static uint16_t sat_u16(float y) {
  if (!(y == y)) return 0;
  if (y >= 65535.0f) return 65535;
  if (y <= 0.0f) return 0;
  return (uint16_t)(uint32_t)y;
}
// max(|x|) from 0 (f32::max ignores NaN, as the comparison below does)
static float max_abs(const float* v, size_t n) {
  float m = 0.0f;
  for (size_t i = 0; i < n; i++) {
    float a = std::fabs(v[i]);
    if (a > m) m = a;
  }
  return m;
}
// storage.rs:236-246: decoded strategy sums
static void load_strat(const Ctx& ctx, const Node& node, size_t len, float* out) {
  if (!ctx.q) {
    std::copy(ctx.strat->data() + node.offset, ctx.strat->data() + node.offset + len, out);
    return;
  }
  const float step = ctx.q->strat_scale[node.slot] / 65535.0f;
  const uint16_t* q = ctx.q->strat.data() + node.offset;
  for (size_t i = 0; i < len; i++) out[i] = (float)q[i] * step;
}
// storage.rs:253-289: f applied to the decoded regrets, then re-encoded with
// the block's new largest |value| as its scale (round half away from zero)
template <typename F> static void update_regrets(const Ctx& ctx, const Node& node, size_t len, F f) {
  if (!ctx.q) {
    f(ctx.regrets->data() + node.offset);
    return;
  }
  float& scale = ctx.q->regret_scale[node.slot];
  int16_t* q = ctx.q->regrets.data() + node.offset;
  std::vector<float> buf(len);
  const float step = scale / 32767.0f;
  for (size_t i = 0; i < len; i++) buf[i] = (float)q[i] * step;
  f(buf.data());
  const float m = max_abs(buf.data(), len);
  scale = m;
  if (m > 0.0f) {
    const float inv = 32767.0f / m;
    for (size_t i = 0; i < len; i++) {
      const float y = buf[i] * inv;
      q[i] = sat_i16(y + std::copysign(0.5f, y));
    }
  } else {
    std::fill(q, q + len, (int16_t)0);
  }
}
// storage.rs:291-325: the same for the strategy sums (u16, values >= 0)
template <typename F> static void update_strat(const Ctx& ctx, const Node& node, size_t len, F f) {
  if (!ctx.q) {
    f(ctx.strat->data() + node.offset);
    return;
  }
  float& scale = ctx.q->strat_scale[node.slot];
  uint16_t* q = ctx.q->strat.data() + node.offset;
  std::vector<float> buf(len);
  const float step = scale / 65535.0f;
  for (size_t i = 0; i < len; i++) buf[i] = (float)q[i] * step;
  f(buf.data());
  const float m = max_abs(buf.data(), len);
  scale = m;
  if (m > 0.0f) {
    const float inv = 65535.0f / m;
    for (size_t i = 0; i < len; i++) {
      const float v = buf[i] > 0.0f ? buf[i] : 0.0f;
      q[i] = sat_u16(v * inv + 0.5f);
    }
  } else {
    std::fill(q, q + len, (uint16_t)0);
  }
}

// One block through the 16-bit codec from zero storage (storage.rs:333-357):
// the stored integers and the new scale.
void encode_block(bool sums, const float* v, size_t n, int16_t* rq, uint16_t* sq, float* scale) {
  Storage16 st;
  st.regrets.assign(n, 0);
  st.strat.assign(n, 0);
  st.regret_scale.assign(1, 0.0f);
  st.strat_scale.assign(1, 0.0f);
  Ctx c{};
  c.q = &st;
  Node nd{};
  nd.offset = 0;
  nd.slot = 0;
  auto copy_in = [&](float* b) { std::copy(v, v + n, b); };
  if (sums) {
    update_strat(c, nd, n, copy_in);
    std::copy(st.strat.begin(), st.strat.end(), sq);
    *scale = st.strat_scale[0];
  } else {
    update_regrets(c, nd, n, copy_in);
    std::copy(st.regrets.begin(), st.regrets.end(), rq);
    *scale = st.regret_scale[0];
  }
}

// cfr.rs:315-382: value of a traverser node from its children (action-major
// cfv); in CFR mode also the regret update. V = float for CFR.
template <typename V> static void own_node_update(const Ctx& ctx, const Node& node, size_t n, const V* cfv, Mode mode, V* out) {
  size_t na = node.num_children;
  if (mode.kind == MBest) {
    for (size_t h = 0; h < n; h++) out[h] = cfv[h];
    for (size_t a = 1; a < na; a++)
      for (size_t h = 0; h < n; h++) { V v = cfv[a * n + h]; out[h] = out[h] > v ? out[h] : (v > out[h] ? v : out[h]); }
    return;
  }
  if (mode.kind == MAvg) {
    std::vector<V> s(na * n);
    std::vector<float> src(na * n);
    load_strat(ctx, node, na * n, src.data());
    for (size_t i = 0; i < na * n; i++) s[i] = (V)src[i];
    normalize_in_place<V>(s.data(), na, n);
    for (size_t h = 0; h < n; h++) out[h] = V(0);
    for (size_t a = 0; a < na; a++)
      for (size_t h = 0; h < n; h++) out[h] += s[a * n + h] * cfv[a * n + h];
    return;
  }
  if constexpr (std::is_same<V, float>::value) {
   update_regrets(ctx, node, na * n, [&](float* r) {
    const Discount& d = mode.d;
    std::vector<float> inv(n);
    for (size_t h = 0; h < n; h++) inv[h] = pos0(r[h]);
    for (size_t a = 1; a < na; a++)
      for (size_t h = 0; h < n; h++) inv[h] += pos0(r[a * n + h]);
    for (size_t h = 0; h < n; h++) out[h] = 0.0f;
    for (size_t a = 0; a < na; a++)
      for (size_t h = 0; h < n; h++) out[h] += pos0(r[a * n + h]) * cfv[a * n + h];
    bool any_zero = false;
    for (size_t h = 0; h < n; h++) { inv[h] = recip(inv[h]); any_zero |= inv[h] == 0.0f; out[h] *= inv[h]; }
    if (any_zero) {
      float mean = 1.0f / (float)na;
      for (size_t h = 0; h < n; h++) {
        if (inv[h] != 0.0f) continue;
        float s = 0.0f;
        for (size_t a = 0; a < na; a++) s += cfv[a * n + h];
        out[h] = s * mean;
      }
    }
    for (size_t a = 0; a < na; a++)
      for (size_t h = 0; h < n; h++) {
        float x = r[a * n + h];
        float disc = x > 0.0f ? d.pos : d.neg;
        r[a * n + h] = x * disc + cfv[a * n + h] - out[h];
      }
   });
  }
}

// The opponent's strategy at an opponent node, with the strategy-sum update
// in CFR mode (cfr.rs:389-414, river.rs:116-144).
template <typename V> static void opponent_strategy(const Ctx& ctx, const Node& node, size_t nq, const V* reach, Mode mode, V* s) {
  size_t na = node.num_children;
  if (mode.kind == MCfr) {
    if constexpr (std::is_same<V, float>::value) {
      // storage.rs:221-234: 16-bit regrets are matched undecoded (as f32
      // integers; the scale cancels)
      if (ctx.q) {
        std::vector<float> r(na * nq);
        const int16_t* q = ctx.q->regrets.data() + node.offset;
        for (size_t i = 0; i < na * nq; i++) r[i] = (float)q[i];
        regret_match_into(r.data(), na, nq, s);
      } else {
        regret_match_into(ctx.regrets->data() + node.offset, na, nq, s);
      }
      float g = mode.d.strat;
      update_strat(ctx, node, na * nq, [&](float* ss) {
        for (size_t a = 0; a < na; a++)
          for (size_t h = 0; h < nq; h++) {
            float& x = ss[a * nq + h];
            if (mode.d.reach_weighted) x = x * g + s[a * nq + h] * reach[h];
            else x = x * g + s[a * nq + h];
          }
      });
    }
  } else {
    std::vector<float> src(na * nq);
    load_strat(ctx, node, na * nq, src.data());
    for (size_t i = 0; i < na * nq; i++) s[i] = (V)src[i];
    normalize_in_place<V>(s, na, nq);
  }
}

template <typename V> static void own_node(const Ctx& ctx, const Node& node, int p, const V* opp, size_t nq, Mode mode, V* out, size_t n) {  // cfr.rs:290-301
  size_t na = node.num_children;
  std::vector<V> cfv(na * (n > 0 ? n : 1));
  for (size_t a = 0; a < na; a++) traverse_t<V>(ctx, node.first_child + a, p, opp, nq, mode, cfv.data() + a * n, n);
  if (n == 0) return;
  own_node_update<V>(ctx, node, n, cfv.data(), mode, out);
}

template <typename V> static void opponent_node(const Ctx& ctx, const Node& node, int p, const V* opp, size_t nq, Mode mode, V* out, size_t n) {  // cfr.rs:384-433
  size_t na = node.num_children;
  std::vector<V> s(na * nq);
  if (nq > 0) {
    opponent_strategy<V>(ctx, node, nq, opp, mode, s.data());
    for (size_t a = 0; a < na; a++)
      for (size_t h = 0; h < nq; h++) s[a * nq + h] *= opp[h];
  }
  for (size_t h = 0; h < n; h++) out[h] = V(0);
  std::vector<V> tmp(n);
  for (size_t a = 0; a < na; a++) {
    const V* ra = s.data() + a * nq;
    if (all_zero(ra, nq)) continue;
    traverse_t<V>(ctx, node.first_child + a, p, ra, nq, mode, tmp.data(), n);
    for (size_t h = 0; h < n; h++) out[h] += tmp[h];
  }
}

template <typename V> static void chance_node(const Ctx& ctx, const Node& node, int p, const V* opp, size_t nq, Mode mode, V* out, size_t n) {  // cfr.rs:435-479
  if (n == 0 || all_zero(opp, nq)) { for (size_t h = 0; h < n; h++) out[h] = V(0); return; }
  int o = p ^ 1;
  V factor = (V)node.chance_factor;
  size_t k = node.num_children;
  std::vector<V> rows(k * n);
  auto child = [&](size_t ci) {
    size_t c = node.first_child + ci;
    int card = ctx.tree->nodes[c].edge_card;
    std::vector<V> r(opp, opp + nq);
    for (size_t h = 0; h < nq; h++) r[h] *= factor;
    for (uint32_t h : ctx.hands->blocked[o][card]) r[h] = V(0);
    V* row = rows.data() + ci * n;
    traverse_t<V>(ctx, c, p, r.data(), nq, mode, row, n);
    for (uint32_t h : ctx.hands->blocked[p][card]) row[h] = V(0);
  };
  // rayon's par_iter over chance children (cfr.rs:448): the children touch
  // disjoint storage, so the result doesn't depend on the schedule.
  if (ctx.parallel && !omp_in_parallel()) {
#pragma omp parallel for schedule(dynamic, 1)
    for (long ci = 0; ci < (long)k; ci++) child((size_t)ci);
  } else {
    for (size_t ci = 0; ci < k; ci++) child(ci);
  }
  std::vector<double> acc(n, 0.0);
  for (size_t ci = 0; ci < k; ci++)
    for (size_t h = 0; h < n; h++) acc[h] += (double)rows[ci * n + h];
  for (auto& pr : ctx.tree->chance[node.slot].iso) {
    const V* row = rows.data() + (size_t)pr.first * n;
    const std::vector<uint16_t>& perm = ctx.iso->perm[p][pr.second];
    for (size_t h = 0; h < n; h++) acc[h] += (double)row[perm[h]];
  }
  for (size_t h = 0; h < n; h++) out[h] = (V)acc[h];
}

// river.rs:66-276: breadth-first river subtree.
template <typename V> static void river_run(const Ctx& ctx, size_t root, int p, const V* opp, size_t nq, Mode mode, V* out, size_t n) {
  const std::vector<Node>& nodes = ctx.tree->nodes;
  struct Item { uint32_t id, first_child, reach; bool dead; };
  std::vector<Item> items;
  items.push_back({(uint32_t)root, 0, 0, all_zero(opp, nq)});
  size_t reach_slots = 1;
  for (size_t k = 0; k < items.size(); k++) {
    const Node& nd = nodes[items[k].id];
    if (is_player(nd.kind)) {
      items[k].first_child = (uint32_t)items.size();
      for (uint32_t c = nd.first_child; c < nd.first_child + nd.num_children; c++) items.push_back({c, 0, 0, false});
      if ((int)nd.kind != p) reach_slots += nd.num_children;
    }
  }
  std::vector<V> reach(reach_slots * nq);
  std::copy(opp, opp + nq, reach.begin());
  size_t next_slot = 1;
  std::vector<V> strat;
  for (size_t k = 0; k < items.size(); k++) {
    Item it = items[k];
    const Node& nd = nodes[it.id];
    if (!is_player(nd.kind)) continue;
    size_t na = nd.num_children, fc = it.first_child;
    if (it.dead) { for (size_t c = fc; c < fc + na; c++) items[c].dead = true; continue; }
    if ((int)nd.kind == p) { for (size_t c = fc; c < fc + na; c++) items[c].reach = it.reach; continue; }
    strat.assign(na * nq, V(0));
    opponent_strategy<V>(ctx, nd, nq, reach.data() + (size_t)it.reach * nq, mode, strat.data());
    size_t parent = (size_t)it.reach * nq;
    for (size_t a = 0; a < na; a++) {
      size_t slot = next_slot++;
      bool any = false;
      for (size_t h = 0; h < nq; h++) {
        V d = reach[parent + h] * strat[a * nq + h];
        reach[slot * nq + h] = d;
        any |= d != V(0);
      }
      items[fc + a].reach = (uint32_t)slot;
      items[fc + a].dead = !any;
    }
  }
  std::vector<V> vals(items.size() * n, V(0));
  if (n == 0) return;
  auto reach_of = [&](uint32_t slot) { return reach.data() + (size_t)slot * nq; };
  std::vector<uint32_t> group_of(reach_slots, NONE);
  struct Group { uint32_t reach, sd, fd; };
  std::vector<Group> groups;
  int64_t board_table = -1;
  for (size_t k = 0; k < items.size(); k++) {
    const Item& it = items[k];
    if (it.dead) continue;
    const Node& nd = nodes[it.id];
    bool is_sd;
    if (nd.kind == KShowdown) is_sd = true;
    else if (is_fold(nd.kind)) is_sd = false;
    else continue;
    if (is_sd) board_table = (int64_t)nd.offset;
    uint32_t g = group_of[it.reach];
    bool fits = g != NONE && (is_sd ? groups[g].sd == NONE : groups[g].fd == NONE);
    if (!fits) {
      group_of[it.reach] = (uint32_t)groups.size();
      groups.push_back({it.reach, NONE, NONE});
      g = (uint32_t)groups.size() - 1;
    }
    if (is_sd) groups[g].sd = (uint32_t)k; else groups[g].fd = (uint32_t)k;
  }
  // showdown groups first; batching into lanes doesn't change any lane's arithmetic
  std::stable_partition(groups.begin(), groups.end(), [](const Group& g) { return g.sd != NONE; });
  size_t split = 0;
  while (split < groups.size() && groups[split].sd != NONE) split++;
  if (board_table >= 0) {
    const StrengthTable& table = (*ctx.tables)[(size_t)board_table];
    for (size_t c0 = 0; c0 < split; c0 += LANES) {
      int k = (int)std::min((size_t)LANES, split - c0);
      const V* reaches[LANES]; double win[LANES], lose[LANES], tie[LANES], fpay[LANES] = {0};
      V* outs[LANES]; V* fouts[LANES] = {nullptr, nullptr, nullptr, nullptr}; int hf[LANES] = {0, 0, 0, 0};
      for (int t = 0; t < k; t++) {
        const Group& g = groups[c0 + t];
        reaches[t] = reach_of(g.reach);
        showdown_payoffs(ctx, nodes[items[g.sd].id], p, win[t], lose[t], tie[t]);
        outs[t] = vals.data() + (size_t)g.sd * n;
        if (g.fd != NONE) {
          hf[t] = 1;
          fpay[t] = fold_payoff(ctx, nodes[items[g.fd].id], p);
          fouts[t] = vals.data() + (size_t)g.fd * n;
        }
      }
      showdown_batch_t<V>(*ctx.hands, table, p, k, reaches, win, lose, tie, outs, hf, fpay, fouts);
    }
  }
  for (size_t c0 = split; c0 < groups.size(); c0 += LANES) {
    int k = (int)std::min((size_t)LANES, groups.size() - c0);
    const V* reaches[LANES]; double pay[LANES]; V* outs[LANES];
    for (int t = 0; t < k; t++) {
      const Group& g = groups[c0 + t];
      reaches[t] = reach_of(g.reach);
      pay[t] = fold_payoff(ctx, nodes[items[g.fd].id], p);
      outs[t] = vals.data() + (size_t)g.fd * n;
    }
    fold_batch_t<V>(*ctx.hands, p, k, reaches, pay, outs);
  }
  for (size_t k = items.size(); k-- > 0;) {
    const Item& it = items[k];
    const Node& nd = nodes[it.id];
    if (!is_player(nd.kind) || it.dead) continue;
    size_t na = nd.num_children, fc = it.first_child;
    V* me = vals.data() + k * n;
    const V* child = vals.data() + fc * n;
    if ((int)nd.kind != p) {
      for (size_t h = 0; h < n; h++) me[h] = V(0);
      for (size_t a = 0; a < na; a++)
        for (size_t h = 0; h < n; h++) me[h] += child[a * n + h];
      continue;
    }
    own_node_update<V>(ctx, nd, n, child, mode, me);
  }
  std::copy(vals.begin(), vals.begin() + n, out);
}

template <typename V> static void traverse_t(const Ctx& ctx, size_t id, int p, const V* opp, size_t nq, Mode mode, V* out, size_t n) {  // cfr.rs:264-288
  const Node& node = ctx.tree->nodes[id];
  if (is_player(node.kind) && all_zero(opp, nq)) { for (size_t h = 0; h < n; h++) out[h] = V(0); return; }
  switch (node.kind) {
    case KFold0: case KFold1: eval_fold_t<V>(*ctx.hands, p, opp, fold_payoff(ctx, node, p), out); return;
    case KShowdown: {
      double w, l, t;
      showdown_payoffs(ctx, node, p, w, l, t);
      eval_showdown_t<V>(*ctx.hands, (*ctx.tables)[node.offset], p, opp, w, l, t, out);
      return;
    }
    case KChance: chance_node<V>(ctx, node, p, opp, nq, mode, out, n); return;
    default: break;
  }
  if (ctx.batch_river && node.street == River) { river_run<V>(ctx, id, p, opp, nq, mode, out, n); return; }
  if ((int)node.kind == p) own_node<V>(ctx, node, p, opp, nq, mode, out, n);
  else opponent_node<V>(ctx, node, p, opp, nq, mode, out, n);
}

void traverse(const Ctx& ctx, size_t id, int p, const float* opp, size_t nq, Mode mode, float* out, size_t n) {
  traverse_t<float>(ctx, id, p, opp, nq, mode, out, n);
}
void traverse_f64(const Ctx& ctx, size_t id, int p, const double* opp, size_t nq, ModeKind mode, double* out, size_t n) {
  Mode m;
  m.kind = mode;
  m.d = {0, 0, 0, false};
  traverse_t<double>(ctx, id, p, opp, nq, m, out, n);
}

// =================================================================== solver (lib.rs)
uint32_t next_check_gap(bool has_prev, uint32_t t0, double e0, uint32_t t, double e, double target, uint32_t max_gap) {  // lib.rs:158-172
  uint32_t min_gap = std::max(max_gap / 4, 1u);
  if (!has_prev) return max_gap;
  if (!(e < e0 && e > target && t > t0 && target > 0.0)) return max_gap;
  double k = std::log(e0 / e) / std::log((double)t / (double)t0);
  double predicted = (double)t * std::pow(e / target, 1.0 / k);
  double gap = std::ceil(predicted - (double)t);
  if (!std::isfinite(gap)) return max_gap;
  double g = std::max(gap, 0.0);
  uint32_t gi = g >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)g;
  return std::min(std::max(gi, min_gap), max_gap);
}

static std::string validate(const SolveConfig& cfg, HandSet& hs) {  // lib.rs:307-329
  size_t n = cfg.board.size();
  if (n < 3 || n > 5) return "invalid board: " + std::to_string(n) + " cards";
  uint64_t set = 0;
  for (int c : cfg.board) {
    if (c < 0 || c >= 52) return "invalid board: bad card";
    set |= card_bit(c);
  }
  if ((size_t)__builtin_popcountll(set) != n) return "invalid board: duplicate cards";
  if (street_from_board((int)n) != cfg.root.street) return "invalid board: board doesn't match the root street";
  hs = make_handset(cfg.ranges, set, cfg.prune, cfg.keep);
  for (int p = 0; p < 2; p++)
    if (hs.len(p) == 0) return "player " + std::to_string(p) + " has an empty range";
  if (hs.num_combinations <= 0.0) return "player 0 has an empty range";
  return "";
}

static TreeParams params_for(const SolveConfig& cfg, const HandSet& hs, const Isomorphism& iso) {  // lib.rs:331-344
  TreeParams P;
  P.root_street = cfg.root.street;
  P.root_board = &cfg.board;
  P.menu = &cfg.menu;
  P.future_menu = cfg.has_future_menu ? &cfg.future_menu : &cfg.menu;
  P.big_blind = cfg.big_blind;
  P.num_hands[0] = hs.len(0);
  P.num_hands[1] = hs.len(1);
  P.rake_rate = cfg.rake_rate;
  P.rake_cap = cfg.rake_cap;
  P.iso = &iso;
  P.chance_mode = cfg.chance_mode;
  P.sample_turn = cfg.sample_turn;
  P.sample_river = cfg.sample_river;
  return P;
}

// lib.rs:346-352: Auto compresses when the f32 tree would exceed 1 GB
static bool storage_compressed(const SolveConfig& cfg, const TreeStats& st) {
  if (cfg.storage_mode == 1) return true;
  if (cfg.storage_mode == 2) return st.memory_bytes_f32() > (1ull << 30);
  return false;
}

std::string estimate(const SolveConfig& cfg, TreeStats& out) {  // lib.rs:355-362
  HandSet hs;
  std::string e = validate(cfg, hs);
  if (!e.empty()) return e;
  Isomorphism iso = make_iso(hs, cfg.isomorphism);
  TreeParams P = params_for(cfg, hs, iso);
  std::vector<int> dealt;
  out = TreeStats();
  StdRng rng(0);  // tree.rs:337-343
  count_player(P, cfg.root, dealt, rng, out);
  out.compressed = storage_compressed(cfg, out);
  return "";
}

std::string make_solver(const SolveConfig& cfg, Solver& s) {  // lib.rs:365-396
  s.config = cfg;
  std::string e = validate(s.config, s.hands);
  if (!e.empty()) return e;
  s.iso = make_iso(s.hands, s.config.isomorphism);
  TreeParams P = params_for(s.config, s.hands, s.iso);
  std::vector<int> dealt;
  s.stats = TreeStats();
  StdRng rng(0);
  count_player(P, s.config.root, dealt, rng, s.stats);
  s.stats.compressed = storage_compressed(s.config, s.stats);
  uint64_t needed = s.stats.memory_bytes();
  if (needed > s.config.max_memory_bytes) {
    char buf[160];
    snprintf(buf, sizeof buf, "game tree needs ~%.2f GB but the limit is %.2f GB", (double)needed / 1e9, (double)s.config.max_memory_bytes / 1e9);
    return buf;
  }
  Builder b(P, s.config.chance_mode == 1 ? s.config.sample_seed : 0);
  b.t.nodes.push_back(placeholder());
  int32_t inv0[2] = {0, 0};
  b.fill_player(0, s.config.root, inv0, dealt);
  s.tree = std::move(b.t);
  s.regrets.clear();
  s.strat.clear();
  s.q = Storage16();
  if (s.stats.compressed) {  // storage.rs:144-156
    s.q.regrets.assign(s.tree.storage_len, 0);
    s.q.strat.assign(s.tree.storage_len, 0);
    s.q.regret_scale.assign(s.tree.num_player_nodes, 0.0f);
    s.q.strat_scale.assign(s.tree.num_player_nodes, 0.0f);
  } else {
    s.regrets.assign(s.tree.storage_len, 0.0f);
    s.strat.assign(s.tree.storage_len, 0.0f);
  }
  s.tables.resize(s.tree.boards.size());
#pragma omp parallel for schedule(dynamic, 16)
  for (long i = 0; i < (long)s.tree.boards.size(); i++) s.tables[i] = make_table(s.hands, s.tree.boards[i]);
  s.iterations = 0;
  s.exploitability_ = std::numeric_limits<double>::infinity();
  return "";
}

Ctx Solver::ctx() {  // lib.rs:444-455
  Ctx c;
  c.tree = &tree; c.regrets = &regrets; c.strat = &strat; c.hands = &hands; c.tables = &tables; c.iso = &iso;
  c.q = stats.compressed ? &q : nullptr;
  c.half_pot = (double)config.root.pot / 2.0;
  c.norm = 1.0 / hands.num_combinations;
  c.batch_river = batch_river;
  c.parallel = parallel;
  return c;
}

void Solver::cpu_iteration() {  // lib.rs:577-585
  Discount d = make_discount(iterations, config.dcfr);
  Ctx c = ctx();
  Mode m{MCfr, d};
  for (int p = 0; p < 2; p++) {
    std::vector<float> out(hands.len(p), 0.0f);
    traverse(c, 0, p, hands.weights[p ^ 1].data(), hands.len(p ^ 1), m, out.data(), out.size());
  }
}

void Solver::root_values(ModeKind mk, double res[2]) {  // cfr.rs:482-490
  Ctx c = ctx();
  Mode m{mk, {0, 0, 0, false}};
  for (int p = 0; p < 2; p++) {
    std::vector<float> v(hands.len(p), 0.0f);
    traverse(c, 0, p, hands.weights[p ^ 1].data(), hands.len(p ^ 1), m, v.data(), v.size());
    double s = 0.0;
    for (size_t i = 0; i < v.size(); i++) s += (double)v[i] * (double)hands.weights[p][i];
    res[p] = s;
  }
}

void Solver::root_values_f64(ModeKind mk, double res[2]) {
  Ctx c = ctx();
  for (int p = 0; p < 2; p++) {
    std::vector<double> opp(hands.weights[p ^ 1].begin(), hands.weights[p ^ 1].end());
    std::vector<double> v(hands.len(p), 0.0);
    traverse_f64(c, 0, p, opp.data(), opp.size(), mk, v.data(), v.size());
    double s = 0.0;
    for (size_t i = 0; i < v.size(); i++) s += v[i] * (double)hands.weights[p][i];
    res[p] = s;
  }
}

double Solver::exploitability() {  // cfr.rs:494-502
  double br[2];
  root_values(MBest, br);
  if (raked()) {
    double ev[2];
    root_values(MAvg, ev);
    return ((br[0] - ev[0]) + (br[1] - ev[1])) / 2.0;
  }
  return (br[0] + br[1]) / 2.0;
}

double Solver::exploitability_f64() {
  double br[2];
  root_values_f64(MBest, br);
  if (raked()) {
    double ev[2];
    root_values_f64(MAvg, ev);
    return ((br[0] - ev[0]) + (br[1] - ev[1])) / 2.0;
  }
  return (br[0] + br[1]) / 2.0;
}

Progress Solver::solve(const SolveOptions& opts, std::vector<Progress>* history) {  // lib.rs:594-642
  double pot = (double)config.root.pot;
  uint32_t check_every = std::max(opts.check_every, 1u);
  Progress pr{iterations, opts.max_iterations, exploitability_, 100.0 * exploitability_ / pot, false};
  double target = opts.target_exploitability_pct * pot / 100.0;
  uint32_t next_check = iterations + check_every;
  bool has_last = false;
  uint32_t lt = 0;
  double le = 0.0;
  while (iterations < opts.max_iterations) {
    iterations += 1;
    cpu_iteration();
    bool last = iterations >= opts.max_iterations;
    if (iterations >= next_check || last) {
      double e = exploitability();
      exploitability_ = e;
      pr.iteration = iterations;
      pr.exploitability = e;
      pr.exploitability_pct = 100.0 * e / pot;
      bool done = pr.exploitability_pct <= opts.target_exploitability_pct || last;
      pr.finished = done;
      if (history) history->push_back(pr);
      if (done) { pr.finished = true; return pr; }
      uint32_t t = iterations;
      next_check = t + next_check_gap(has_last, lt, le, t, e, target, check_every);
      has_last = true; lt = t; le = e;
    }
  }
  pr.finished = true;
  return pr;
}

std::vector<float> Solver::average_strategy(size_t id, size_t n) {  // lib.rs:735-743
  const Node& node = tree.nodes[id];
  std::vector<float> s(node.num_children * n, 0.0f);
  if (n > 0) {
    Ctx c = const_cast<Solver*>(this)->ctx();
    orc::load_strat(c, node, s.size(), s.data());
    normalize_in_place<float>(s.data(), node.num_children, n);
  }
  return s;
}

std::string Solver::walk(const std::vector<Step>& path, size_t& id, std::vector<float> reach[2], std::vector<uint32_t> to_canon[2]) {  // lib.rs:746-825
  id = 0;
  reach[0] = hands.weights[0];
  reach[1] = hands.weights[1];
  for (int x = 0; x < 2; x++) { to_canon[x].resize(hands.len(x)); for (size_t i = 0; i < hands.len(x); i++) to_canon[x][i] = (uint32_t)i; }
  int suit_map[4] = {0, 1, 2, 3};
  for (size_t depth = 0; depth < path.size(); depth++) {
    const Node node = tree.nodes[id];
    const Step& st = path[depth];
    if (is_player(node.kind) && st.kind == StepAction) {
      std::vector<Action> acts = tree.actions(id);
      size_t ai = acts.size();
      for (size_t i = 0; i < acts.size(); i++) if (acts[i] == st.action) { ai = i; break; }
      if (ai == acts.size()) return "bad path: action not available at depth " + std::to_string(depth);
      int q = (int)node.kind;
      size_t n = hands.len(q);
      std::vector<float> s = average_strategy(id, n);
      for (size_t i = 0; i < n; i++) reach[q][i] *= s[ai * n + i];
      id = node.first_child + ai;
    } else if (node.kind == KChance && st.kind == StepCard) {
      int c = st.card;
      int cc = rank_of(c) * 4 + suit_map[suit_of(c)];
      int canonical = -1;
      for (uint32_t ch = node.first_child; ch < node.first_child + node.num_children; ch++)
        if (tree.nodes[ch].edge_card == cc) { id = ch; canonical = cc; break; }
      if (canonical < 0) {
        const ChanceInfo& info = tree.chance[node.slot];
        size_t j = info.iso_cards.size();
        for (size_t k = 0; k < info.iso_cards.size(); k++) if (info.iso_cards[k] == cc) { j = k; break; }
        if (j == info.iso_cards.size()) return "bad path: card not in the tree at depth " + std::to_string(depth);
        int ci = info.iso[j].first, pair = info.iso[j].second;
        int a = PAIRS[pair][0], b = PAIRS[pair][1];
        for (int& s : suit_map) s = s == a ? b : s == b ? a : s;
        for (int x = 0; x < 2; x++) {
          const std::vector<uint16_t>& perm = iso.perm[x][pair];
          std::vector<float> r2(reach[x].size(), 0.0f);
          for (size_t h = 0; h < reach[x].size(); h++) r2[perm[h]] = reach[x][h];
          reach[x] = r2;
          for (uint32_t& t : to_canon[x]) t = perm[t];
        }
        id = node.first_child + ci;
        canonical = tree.nodes[id].edge_card;
      }
      for (int x = 0; x < 2; x++)
        for (uint32_t k : hands.blocked[x][canonical]) reach[x][k] = 0.0f;
    } else {
      return "bad path: step doesn't match node kind at depth " + std::to_string(depth);
    }
  }
  return "";
}

std::string Solver::query(const std::vector<Step>& path, NodeInfo& info) {  // lib.rs:842-899
  size_t id;
  std::vector<float> reach[2];
  std::vector<uint32_t> to_canon[2];
  std::string e = walk(path, id, reach, to_canon);
  if (!e.empty()) return e;
  const Node node = tree.nodes[id];
  if (!is_player(node.kind)) return "bad path: not a decision node";
  int q = (int)node.kind, o = q ^ 1;
  size_t n = hands.len(q), na = node.num_children;
  Ctx c = ctx();
  std::vector<double> compat = hands.compatible_reach(q, reach[o].data());
  std::vector<std::vector<float>> child_vals(na, std::vector<float>(n, 0.0f));
  Mode m{MAvg, {0, 0, 0, false}};
#pragma omp parallel for schedule(dynamic, 1) if (parallel)
  for (long a = 0; a < (long)na; a++)
    traverse(c, node.first_child + (size_t)a, q, reach[o].data(), reach[o].size(), m, child_vals[a].data(), n);
  double offset = c.half_pot + (double)node.invest[q];
  std::vector<float> strat_c = average_strategy(id, n);
  info.player = q;
  info.actions = tree.actions(id);
  info.combos = hands.combos[q];
  info.strategy.assign(na * n, 0.0f);
  info.ev.assign(na * n, 0.0f);
  info.reach.assign(n, 0.0f);
  for (size_t h = 0; h < n; h++) {
    size_t hc = to_canon[q][h];
    info.reach[h] = reach[q][hc];
    double denom = compat[hc] * c.norm;
    for (size_t a = 0; a < na; a++) {
      info.strategy[a * n + h] = strat_c[a * n + hc];
      info.ev[a * n + h] = denom > 0.0 ? (float)((double)child_vals[a][hc] / denom + offset) : 0.0f;
    }
  }
  for (int p = 0; p < 2; p++) {
    double root = 0.0, now = 0.0;
    for (float x : hands.weights[p]) root += (double)x;
    for (float x : reach[p]) now += (double)x;
    info.reach_fraction[p] = root > 0.0 ? now / root : 0.0;
  }
  info.pot = config.root.pot + node.invest[0] + node.invest[1];
  info.invest[0] = node.invest[0];
  info.invest[1] = node.invest[1];
  return "";
}

std::string Solver::ranges_after(const std::vector<Step>& path, std::vector<float> out[2]) {  // lib.rs:903-912
  size_t id;
  std::vector<float> reach[2];
  std::vector<uint32_t> to_canon[2];
  std::string e = walk(path, id, reach, to_canon);
  if (!e.empty()) return e;
  for (int p = 0; p < 2; p++) {
    out[p].assign(1326, 0.0f);
    for (size_t h = 0; h < hands.len(p); h++) out[p][hands.combos[p][h]] = reach[p][to_canon[p][h]];
  }
  return "";
}

}  // namespace orc
