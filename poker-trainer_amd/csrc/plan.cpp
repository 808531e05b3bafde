// Plan builder and validator (the host-side checks that every index the
// kernels follow is in range, as crates/solver/src/gpu/plan.rs:466-699 does
// for the reference's CUDA plan).
#include "plan.hpp"

#include <algorithm>
#include <cstdio>
#include <deque>

namespace pts {

static std::string fail(const char* what, size_t i) {
  char b[160];
  std::snprintf(b, sizeof b, "invalid plan input: %s (at %zu)", what, i);
  return b;
}

std::string build_plan(const PlanInput& in, Plan& P) {
  const size_t N = in.num_nodes;
  if (N == 0) return fail("empty tree", 0);
  if (in.num_hands[0] == 0 || in.num_hands[1] == 0) return fail("empty hand list", 0);
  if (in.num_hands[0] > 1326 || in.num_hands[1] > 1326) return fail("more than 1326 hands", 0);
  if (!(in.num_combinations > 0.0)) return fail("no compatible hand pairs", 0);
  if (N >= 0xFFFFFFFFull) return fail("tree too large for 32-bit node ids", N);
  P.num_nodes = N;
  P.n[0] = in.num_hands[0];
  P.n[1] = in.num_hands[1];
  P.half_pot = in.half_pot;
  P.norm = 1.0 / in.num_combinations;
  P.storage_len = in.storage_len;
  P.num_pslots = in.pslot ? in.num_pslots : 0;
  P.batch_river = in.batch_river;
  P.num_boards = in.num_boards;

  // ---- hands
  for (int q = 0; q < 2; q++) {
    uint32_t n = in.num_hands[q];
    P.cards[q].assign(in.hand_cards[q], in.hand_cards[q] + 2 * n);
    P.weights[q].assign(in.hand_weights[q], in.hand_weights[q] + n);
    P.same[q].assign(in.hand_same[q], in.hand_same[q] + n);
    P.iso_perm[q].clear();
    if (in.iso_refs && in.iso_perm[q]) {
      P.iso_perm[q].assign(in.iso_perm[q], in.iso_perm[q] + 6 * (size_t)n);
      for (uint16_t x : P.iso_perm[q])
        if (x >= n) return fail("bad isomorphism permutation", x);
    }
    for (uint32_t i = 0; i < n; i++) {
      uint8_t a = P.cards[q][2 * i], b = P.cards[q][2 * i + 1];
      if (a >= 52 || b >= 52 || a == b) return fail("bad hand cards", i);
      if (P.same[q][i] != kNone && P.same[q][i] >= in.num_hands[q ^ 1]) return fail("bad same-combo index", i);
      if (!(P.weights[q][i] > 0.0f)) return fail("non-positive hand weight", i);
    }
    P.lookup[q].assign(1326, kNone);
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t a = P.cards[q][2 * i], b = P.cards[q][2 * i + 1];
      const uint32_t hi = std::max(a, b), lo = std::min(a, b);
      P.lookup[q][hi * (hi - 1) / 2 + lo] = i;
    }
    P.bl_off[q].assign(53, 0);
    std::vector<std::vector<uint16_t>> lists(52);
    for (uint32_t i = 0; i < n; i++) {
      lists[P.cards[q][2 * i]].push_back((uint16_t)i);
      lists[P.cards[q][2 * i + 1]].push_back((uint16_t)i);
    }
    P.bl_idx[q].clear();
    for (int c = 0; c < 52; c++) {
      P.bl_off[q][c] = (uint32_t)P.bl_idx[q].size();
      P.bl_idx[q].insert(P.bl_idx[q].end(), lists[c].begin(), lists[c].end());
    }
    P.bl_off[q][52] = (uint32_t)P.bl_idx[q].size();
  }

  // ---- nodes
  P.kind.assign(in.kind, in.kind + N);
  P.na.assign(N, 0);
  P.first_child.assign(in.first_child, in.first_child + N);
  P.offset.assign(in.offset, in.offset + N);
  P.invest.assign(in.invest, in.invest + 2 * N);
  P.rake.assign(in.rake, in.rake + N);
  P.factor.assign(in.chance_factor, in.chance_factor + N);
  P.card.assign(in.card, in.card + N);
  std::vector<uint32_t> parent(N, kNone);
  if (!is_player(P.kind[0])) return fail("root is not a decision node", 0);
  for (size_t i = 0; i < N; i++) {
    uint8_t k = P.kind[i];
    if (k > kLeaf) return fail("bad node kind", i);
    if (k == kLeaf && !(P.factor[i] > 0.0f)) return fail("bad leaf factor", i);
    uint32_t nc = in.num_children[i];
    if (is_terminal(k)) {
      if (nc != 0) return fail("terminal with children", i);
      if (k == kShowdown && P.offset[i] >= in.num_boards) return fail("showdown board out of range", i);
      continue;
    }
    if (nc == 0 || nc > (is_player(k) ? 16u : 255u)) return fail("bad child count", i);
    P.na[i] = (uint8_t)nc;
    uint64_t f = P.first_child[i];
    if (f <= i || f + nc > N) return fail("children out of range", i);
    for (uint32_t a = 0; a < nc; a++) {
      if (parent[f + a] != kNone) return fail("node with two parents", f + a);
      parent[f + a] = (uint32_t)i;
    }
    if (is_player(k)) {
      uint64_t cells = (uint64_t)nc * in.num_hands[k];
      if (P.offset[i] + cells > in.storage_len) return fail("storage block out of range", i);
      if (in.pslot && in.pslot[i] >= in.num_pslots) return fail("player slot out of range", i);
    } else {
      if (!(P.factor[i] > 0.0f)) return fail("bad chance factor", i);
      for (uint32_t a = 0; a < nc; a++)
        if (P.card[f + a] < 0 || P.card[f + a] >= 52) return fail("chance child without a card", f + a);
    }
  }
  for (size_t i = 1; i < N; i++)
    if (parent[i] == kNone) return fail("unreachable node", i);
  P.subtree_end.assign(N, 0);
  for (size_t i = N; i-- > 0;) {
    uint32_t e = (uint32_t)i + 1;
    if (!is_terminal(P.kind[i]))
      for (uint32_t a = 0; a < P.na[i]; a++) e = std::max(e, P.subtree_end[P.first_child[i] + a]);
    P.subtree_end[i] = e;
  }
  for (uint32_t b = 0; b < in.num_boards; b++) {
    uint64_t s = 0;
    for (int k = 0; k < 5; k++) {
      int c = in.boards[5 * b + k];
      if (c < 0 || c >= 52 || (s >> c & 1)) return fail("bad board", b);
      s |= 1ull << c;
    }
  }

  // ---- segments (breadth-first members; chance children start new segments)
  struct Seg { uint32_t root, parent, depth; std::vector<uint32_t> mem; };
  std::vector<Seg> segs;
  std::deque<std::pair<uint32_t, uint32_t>> roots;  // (root, depth)
  roots.push_back({0, 0});
  P.local.assign(N, 0);
  while (!roots.empty()) {
    auto [r, d] = roots.front();
    roots.pop_front();
    Seg s{r, r == 0 ? kNone : parent[r], d, {r}};
    for (size_t m = 0; m < s.mem.size(); m++) {
      uint32_t id = s.mem[m];
      uint8_t k = P.kind[id];
      if (is_player(k)) {
        for (uint32_t a = 0; a < P.na[id]; a++) s.mem.push_back(P.first_child[id] + a);
      } else if (k == kChance) {
        for (uint32_t a = 0; a < P.na[id]; a++) roots.push_back({P.first_child[id] + a, d + 1});
      }
    }
    if (s.mem.size() > 0xFFFF) return fail("segment too large", r);
    for (size_t m = 0; m < s.mem.size(); m++) P.local[s.mem[m]] = (uint16_t)m;
    segs.push_back(std::move(s));
    if (segs.size() >= 0xFFFFFFFFull) return fail("too many segments", r);
  }
  std::stable_sort(segs.begin(), segs.end(), [](const Seg& a, const Seg& b) { return a.depth < b.depth; });
  uint32_t maxd = segs.back().depth;
  P.depth_off.assign(maxd + 2, 0);
  P.members.clear();
  P.max_members = 0;
  for (size_t i = 0; i < segs.size(); i++) {
    P.depth_off[segs[i].depth + 1] = (uint32_t)(i + 1);
    P.seg_root.push_back(segs[i].root);
    P.seg_parent.push_back(segs[i].parent);
    P.seg_mem_off.push_back((uint32_t)P.members.size());
    P.seg_mem_cnt.push_back((uint32_t)segs[i].mem.size());
    P.members.insert(P.members.end(), segs[i].mem.begin(), segs[i].mem.end());
    P.max_members = std::max(P.max_members, (uint32_t)segs[i].mem.size());
  }
  for (uint32_t d = 1; d <= maxd + 1; d++) P.depth_off[d] = std::max(P.depth_off[d], P.depth_off[d - 1]);
  P.seg_of.assign(N, kNone);
  for (size_t i = 0; i < segs.size(); i++)
    for (uint32_t m = 0; m < P.seg_mem_cnt[i]; m++) P.seg_of[P.members[P.seg_mem_off[i] + m]] = (uint32_t)i;

  // ---- chunks: the subtrees rooted at depth-1 segments (the root street's
  // chance children). A traversal runs them in groups sized to the scratch
  // budget, so every reach slot and value row used only inside one chunk is
  // numbered contiguously per chunk; rows read across chunks (the top
  // segment's, the chunk roots') come first and stay resident.
  const uint32_t D = P.depths();
  P.num_chunks = D >= 2 ? P.depth_off[2] - P.depth_off[1] : 0;
  const uint32_t C = P.num_chunks;
  std::vector<uint32_t> seg_chunk(P.seg_root.size(), kNone);
  for (uint32_t g = 0; g < (uint32_t)P.seg_root.size(); g++) {
    if (g < P.depth_off[1]) continue;
    if (g < P.depth_off[std::min(2u, D)]) seg_chunk[g] = g - P.depth_off[1];
    else seg_chunk[g] = seg_chunk[P.seg_of[P.seg_parent[g]]];
  }
  P.chunk_of.assign(N, kNone);
  for (size_t i = 0; i < N; i++) P.chunk_of[i] = seg_chunk[P.seg_of[i]];
  // per depth d >= 1 the segments of chunk c are [seg_coff[d][c], seg_coff[d][c+1])
  P.seg_coff.assign((size_t)D * (C + 1), 0);
  for (uint32_t d = 1; d < D; d++) {
    uint32_t* o = P.seg_coff.data() + (size_t)d * (C + 1);
    uint32_t g = P.depth_off[d];
    for (uint32_t c = 0; c <= C; c++) {
      while (g < P.depth_off[d + 1] && seg_chunk[g] < c) g++;
      o[c] = g;
    }
    for (uint32_t h = P.depth_off[d]; h + 1 < P.depth_off[d + 1]; h++)
      if (seg_chunk[h] > seg_chunk[h + 1]) return fail("segments not ordered by chunk", h);
  }
  // nodes of each chunk in arena order (parents first)
  std::vector<uint32_t> chunk_nodes_off(C + 1, 0), chunk_nodes;
  {
    std::vector<uint32_t> cnt(C + 1, 0);
    for (size_t i = 0; i < N; i++)
      if (P.chunk_of[i] != kNone) cnt[P.chunk_of[i] + 1]++;
    for (uint32_t c = 0; c < C; c++) cnt[c + 1] += cnt[c];
    chunk_nodes_off = cnt;
    chunk_nodes.resize(cnt[C]);
    for (size_t i = 0; i < N; i++)
      if (P.chunk_of[i] != kNone) chunk_nodes[cnt[P.chunk_of[i]]++] = (uint32_t)i;
  }

  // ---- reach slots and alive sources, per traverser: the top segment's
  // children first (chunk roots included), then chunk by chunk
  P.rslot.assign(2 * N, kNone);
  P.alive_src.assign(2 * N, 0);
  for (int p = 0; p < 2; p++) {
    uint32_t* rs = P.rslot.data() + p * N;
    uint16_t* src = P.alive_src.data() + p * N;
    uint64_t next = 1;
    rs[0] = 0;
    src[0] = 0;
    auto give = [&](size_t i) {
      uint8_t k = P.kind[i];
      if (is_terminal(k)) return;
      for (uint32_t a = 0; a < P.na[i]; a++) {
        uint32_t c = P.first_child[i] + a;
        bool fresh = k == kChance || (int)k == (p ^ 1);
        rs[c] = fresh ? (uint32_t)next++ : rs[i];
        src[c] = fresh ? P.local[c] : src[i];
      }
    };
    for (size_t i = 0; i < N; i++)
      if (P.chunk_of[i] == kNone) give(i);
    P.slot_s0[p] = (uint32_t)next;
    P.slot_coff[p].assign(C + 1, 0);
    for (uint32_t c = 0; c < C; c++) {
      P.slot_coff[p][c] = (uint32_t)next;
      for (uint32_t j = chunk_nodes_off[c]; j < chunk_nodes_off[c + 1]; j++) give(chunk_nodes[j]);
    }
    P.slot_coff[p][C] = (uint32_t)next;
    if (next >= 0xFFFFFFFFull) return fail("too many reach slots", next);
    P.num_slots[p] = (uint32_t)next;
  }

  // ---- value rows: the root and chunk roots, the top segment's chance
  // nodes and terminals; then per chunk its deeper segment roots, chance
  // nodes and terminals
  P.vslot.assign(N, kNone);
  P.tflags.assign(N, 0);
  P.tboard.assign(N, kNone);
  uint64_t nv = 0;
  std::vector<uint32_t> chance_vslot(N, kNone);
  for (uint32_t g = 0; g < P.depth_off[std::min(2u, D)]; g++) P.vslot[P.seg_root[g]] = (uint32_t)nv++;
  for (uint32_t g = 0; g < P.depth_off[1]; g++)
    for (uint32_t m = 0; m < P.seg_mem_cnt[g]; m++) {
      uint32_t id = P.members[P.seg_mem_off[g] + m];
      if (P.kind[id] == kChance && P.vslot[id] == kNone) chance_vslot[id] = (uint32_t)nv++;
    }
  P.terminals.clear();
  P.leaves.clear();
  for (size_t i = 0; i < N; i++)
    if (P.kind[i] == kLeaf) {
      if (P.chunk_of[i] != kNone) return fail("leaf below a chance node", i);
      LeafRec l{};
      l.node = (uint32_t)i;
      l.vslot = (uint32_t)nv++;
      l.slot[0] = P.rslot[i];
      l.slot[1] = P.rslot[N + i];
      l.pot = (uint32_t)P.offset[i];
      const int64_t s0 = in.root_stack[0] - P.invest[2 * i], s1 = in.root_stack[1] - P.invest[2 * i + 1];
      l.stack = (uint32_t)std::max<int64_t>(0, std::min(s0, s1));
      l.factor = P.factor[i];
      l.local = P.local[i];
      l.spot = 0;
      P.vslot[i] = l.vslot;
      P.leaves.push_back(l);
    }
  for (size_t i = 0; i < N; i++)
    if (P.chunk_of[i] == kNone && is_terminal(P.kind[i]) && P.kind[i] != kLeaf) {
      P.terminals.push_back((uint32_t)i);
      if (P.vslot[i] == kNone) P.vslot[i] = (uint32_t)nv++;  // a chance child keeps its segment-root row
    }
  P.val_v0 = (uint32_t)nv;
  P.val_coff.assign(C + 1, 0);
  P.term_coff.assign(C + 1, 0);
  for (uint32_t c = 0; c < C; c++) {
    P.val_coff[c] = (uint32_t)nv;
    P.term_coff[c] = (uint32_t)P.terminals.size();
    for (uint32_t d = 2; d < D; d++) {
      const uint32_t* o = P.seg_coff.data() + (size_t)d * (C + 1);
      for (uint32_t g = o[c]; g < o[c + 1]; g++) P.vslot[P.seg_root[g]] = (uint32_t)nv++;
    }
    for (uint32_t j = chunk_nodes_off[c]; j < chunk_nodes_off[c + 1]; j++) {
      uint32_t id = chunk_nodes[j];
      if (P.kind[id] == kChance && P.vslot[id] == kNone) chance_vslot[id] = (uint32_t)nv++;
    }
    for (uint32_t j = chunk_nodes_off[c]; j < chunk_nodes_off[c + 1]; j++) {
      uint32_t id = chunk_nodes[j];
      if (!is_terminal(P.kind[id])) continue;
      P.terminals.push_back(id);
      if (P.vslot[id] == kNone) P.vslot[id] = (uint32_t)nv++;
    }
  }
  P.val_coff[C] = (uint32_t)nv;
  P.term_coff[C] = (uint32_t)P.terminals.size();
  if (nv >= 0xFFFFFFFFull) return fail("too many value rows", nv);
  for (uint32_t id : P.terminals) {
    uint32_t par = parent[id];
    bool batched = in.batch_river && is_player(P.kind[par]) && in.street[par] == kRiver;
    if (batched) P.tflags[id] |= kTermBatched;
    if (P.kind[id] == kShowdown) P.tboard[id] = (uint32_t)P.offset[id];
  }
  // A fold and a showdown under one river node of the traverser share its
  // reach and one strength sweep (crates/solver/src/river.rs:163-213).
  for (size_t i = 0; i < N; i++) {
    if (!(P.tflags[i] & kTermBatched) || (P.kind[i] != kFold0 && P.kind[i] != kFold1)) continue;
    uint32_t par = parent[i];
    for (uint32_t a = 0; a < P.na[par]; a++) {
      uint32_t s = P.first_child[par] + a;
      if (P.kind[s] != kShowdown) continue;
      P.tflags[i] |= P.kind[par] == kP0 ? kTermPaired0 : kTermPaired1;
      P.tboard[i] = (uint32_t)P.offset[s];
      break;
    }
  }
  P.num_vslots = (uint32_t)nv;

  // ---- member and terminal records
  const size_t S = P.seg_root.size();
  P.seg_pslot.assign(2 * S, kNone);
  P.seg_factor.assign(S, 0.0f);
  for (size_t g = 0; g < S; g++) {
    uint32_t par = P.seg_parent[g];
    if (par == kNone) continue;
    P.seg_pslot[g] = P.rslot[par];
    P.seg_pslot[S + g] = P.rslot[N + par];
    P.seg_factor[g] = P.factor[par];
  }
  P.mrec.assign(P.members.size(), MemRec{});
  for (size_t m = 0; m < P.members.size(); m++) {
    uint32_t id = P.members[m];
    MemRec& r = P.mrec[m];
    r.node = id;
    r.kind = P.kind[id];
    r.na = P.na[id];
    r.card = P.card[id];
    r.tflags = P.tflags[id];
    r.vslot = P.vslot[id];
    for (int p = 0; p < 2; p++) {
      r.slot[p] = P.rslot[(size_t)p * N + id];
      r.src[p] = P.alive_src[(size_t)p * N + id];
      r.cslot[p] = kNone;
    }
    if (is_terminal(r.kind)) continue;
    uint32_t fc = P.first_child[id];
    for (int p = 0; p < 2; p++) {
      const uint32_t* rs = P.rslot.data() + (size_t)p * N;
      bool fresh = r.kind == kChance || (int)r.kind == (p ^ 1);
      if (fresh) {
        r.cslot[p] = rs[fc];
        for (uint32_t a = 0; a < r.na; a++)
          if (rs[fc + a] != rs[fc] + a) return fail("child reach slots not consecutive", id);
      }
    }
    if (r.kind == kChance) {
      r.offset = fc;
      for (uint32_t a = 0; a < r.na; a++)
        if (P.vslot[fc + a] != P.vslot[fc] + a) return fail("chance children value rows not consecutive", id);
      // a chance node that is itself a segment root keeps that row; the
      // chance kernel writes the node's value there
      if (r.vslot == kNone) r.vslot = chance_vslot[id];
    } else {
      r.offset = P.offset[id];
      r.pslot = in.pslot ? in.pslot[id] : 0;
      r.clocal = P.local[fc];
      for (uint32_t a = 0; a < r.na; a++)
        if (P.local[fc + a] != P.local[fc] + a || P.seg_of[fc + a] != P.seg_of[id]) return fail("children not consecutive members", id);
    }
  }
  // chance nodes by segment depth
  {
    P.chance_off.assign(D + 1, 0);
    P.crec.clear();
    P.chance_child_vslot.clear();
    P.ciso_off.assign(1, 0);
    P.ciso.clear();
    P.ch_coff.assign((size_t)D * (C + 1), 0);
    for (uint32_t d = 0; d < D; d++) {
      P.chance_off[d] = (uint32_t)P.crec.size();
      uint32_t* co = P.ch_coff.data() + (size_t)d * (C + 1);
      uint32_t c = 0;
      for (uint32_t g = P.depth_off[d]; g < P.depth_off[d + 1]; g++) {
        if (d >= 1)
          while (c <= seg_chunk[g]) co[c++] = (uint32_t)P.crec.size();
        for (uint32_t m = 0; m < P.seg_mem_cnt[g]; m++) {
          const MemRec& r = P.mrec[P.seg_mem_off[g] + m];
          if (r.kind != kChance) continue;
          P.crec.push_back(r);
          P.chance_child_vslot.push_back(P.vslot[r.offset]);
          if (in.iso_refs)
            for (uint32_t k = 0; k < in.iso_cnt[r.node]; k++) {
              const IsoRef& e = in.iso_refs[in.iso_first[r.node] + k];
              if (e.child >= r.na || e.pair >= 6 || e.card < 0 || e.card >= 52 || P.iso_perm[0].empty())
                return fail("bad isomorphism reference", r.node);
              P.ciso.push_back(e);
            }
          P.ciso_off.push_back((uint32_t)P.ciso.size());
        }
      }
      if (d >= 1)
        while (c <= C) co[c++] = (uint32_t)P.crec.size();
    }
    P.chance_off[D] = (uint32_t)P.crec.size();
  }
  // Folds inside a river subtree get the subtree's board, so the fast
  // terminal kernel can sum them in that board's strength order (the exact
  // kernel reads the board of paired folds only).
  {
    std::vector<uint32_t> seg_board(P.seg_root.size(), kNone);
    for (uint32_t id : P.terminals)
      if (P.kind[id] == kShowdown && seg_board[P.seg_of[id]] == kNone) seg_board[P.seg_of[id]] = (uint32_t)P.offset[id];
    for (uint32_t id : P.terminals)
      if (P.kind[id] != kShowdown && P.tboard[id] == kNone && in.street[parent[id]] == kRiver)
        P.tboard[id] = seg_board[P.seg_of[id]];
  }
  // Within the top segment's terminals and within each chunk's, the ones on
  // one river board together (in tree order otherwise), so the fast terminal
  // kernel's runs load a board's tables once for up to kBatchMax terminals:
  // all-in runouts deal one showdown per segment, and without this each
  // would be a run of its own. Siblings on one board stay adjacent.
  {
    auto by_board = [&](uint32_t a, uint32_t b) {
      std::stable_sort(P.terminals.begin() + a, P.terminals.begin() + b,
                       [&](uint32_t x, uint32_t y) { return P.tboard[x] < P.tboard[y]; });
    };
    by_board(0, C ? P.term_coff[0] : (uint32_t)P.terminals.size());
    for (uint32_t c = 0; c < C; c++) by_board(P.term_coff[c], P.term_coff[c + 1]);
  }
  P.trec.assign(P.terminals.size(), TermRec{});
  for (size_t t = 0; t < P.terminals.size(); t++) {
    uint32_t id = P.terminals[t];
    TermRec& r = P.trec[t];
    r.node = id;
    r.slot[0] = P.rslot[id];
    r.slot[1] = P.rslot[N + id];
    r.vslot = P.vslot[id];
    r.board = P.tboard[id];
    r.invest[0] = P.invest[2 * id];
    r.invest[1] = P.invest[2 * id + 1];
    r.rake = P.rake[id];
    r.kind = P.kind[id];
    r.tflags = P.tflags[id];
  }

  // ---- terminal runs (per chunk, consecutive terminals with one board)
  build_runs(P, default_run_length(P));

  // ---- the one spot's root
  P.num_spots = 1;
  P.parts.clear();
  P.spot_root_node.assign(1, 0);
  P.spot_root_vslot.assign(1, P.vslot[0]);
  P.spot_root_slot[0].assign(1, P.rslot[0]);
  P.spot_root_slot[1].assign(1, P.rslot[N]);

  // ---- strength tables (crates/solver/src/hands.rs:128-159)
  uint32_t B = in.num_boards;
  for (int q = 0; q < 2; q++) {
    P.ent_off[q].assign(B, 0);
    P.ent_cnt[q].assign(B, 0);
    P.cl_base[q].assign(B, 0);
    P.cl_off[q].assign((size_t)B * 53, 0);
    P.hand_ent[q].assign((size_t)B * P.n[q], 0xFFFF);
    P.ent_idx[q].clear();
    P.ent_str[q].clear();
    P.cl_idx[q].clear();
  }
  std::vector<std::vector<uint32_t>> cl_str[2];
  for (uint32_t b = 0; b < B; b++) {
    uint64_t lanes = 0, bits = 0;
    for (int k = 0; k < 5; k++) { lanes |= lane_bit(in.boards[5 * b + k]); bits |= 1ull << in.boards[5 * b + k]; }
    for (int q = 0; q < 2; q++) {
      std::vector<std::pair<uint32_t, uint16_t>> e;
      for (uint32_t i = 0; i < P.n[q]; i++) {
        int a = P.cards[q][2 * i], c = P.cards[q][2 * i + 1];
        if ((bits >> a & 1) || (bits >> c & 1)) continue;
        e.push_back({hand_strength(lanes | lane_bit(a) | lane_bit(c)), (uint16_t)i});
      }
      std::sort(e.begin(), e.end());  // (strength, hand index): the reference's Entry order
      P.ent_off[q][b] = (uint32_t)P.ent_idx[q].size();
      P.ent_cnt[q][b] = (uint32_t)e.size();
      P.max_entries = std::max(P.max_entries, (uint32_t)e.size());
      std::vector<std::vector<std::pair<uint32_t, uint16_t>>> per(52);
      for (size_t j = 0; j < e.size(); j++) {
        P.ent_idx[q].push_back(e[j].second);
        P.ent_str[q].push_back(e[j].first);
        P.hand_ent[q][(size_t)b * P.n[q] + e[j].second] = (uint16_t)j;
        per[P.cards[q][2 * e[j].second]].push_back(e[j]);
        per[P.cards[q][2 * e[j].second + 1]].push_back(e[j]);
      }
      P.cl_base[q][b] = (uint32_t)P.cl_idx[q].size();
      cl_str[q].assign(52, {});
      uint32_t pos = 0;
      for (int c = 0; c < 52; c++) {
        P.cl_off[q][(size_t)b * 53 + c] = (uint16_t)pos;
        for (auto& x : per[c]) { P.cl_idx[q].push_back(x.second); cl_str[q][c].push_back(x.first); }
        pos += (uint32_t)per[c].size();
      }
      P.cl_off[q][(size_t)b * 53 + 52] = (uint16_t)pos;
    }
    // counts of the other player's entries below / at each of q's strengths
    for (int q = 0; q < 2; q++) {
      int o = q ^ 1;
      const uint32_t* os = P.ent_str[o].data() + P.ent_off[o][b];
      uint32_t om = P.ent_cnt[o][b];
      if (q == 0 && b == 0) {
        P.my_cnt[0].clear();
        P.my_cnt[1].clear();
      }
      for (uint32_t j = 0; j < P.ent_cnt[q][b]; j++) {
        uint32_t s = P.ent_str[q][P.ent_off[q][b] + j];
        uint16_t h = P.ent_idx[q][P.ent_off[q][b] + j];
        uint16_t lt = (uint16_t)(std::lower_bound(os, os + om, s) - os);
        uint16_t le = (uint16_t)(std::upper_bound(os, os + om, s) - os);
        P.my_cnt[q].push_back(lt);
        P.my_cnt[q].push_back(le);
        for (int k = 0; k < 2; k++) {
          const std::vector<uint32_t>& cs = cl_str[o][P.cards[q][2 * h + k]];
          P.my_cnt[q].push_back((uint16_t)(std::lower_bound(cs.begin(), cs.end(), s) - cs.begin()));
          P.my_cnt[q].push_back((uint16_t)(std::upper_bound(cs.begin(), cs.end(), s) - cs.begin()));
        }
      }
    }
  }
  for (int q = 0; q < 2; q++) {
    P.hand_cnt[q].assign((size_t)B * P.n[q] * 8, 0);
    for (uint32_t b = 0; b < B; b++)
      for (uint32_t h = 0; h < P.n[q]; h++) {
        uint16_t* d = P.hand_cnt[q].data() + ((size_t)b * P.n[q] + h) * 8;
        uint16_t e = P.hand_ent[q][(size_t)b * P.n[q] + h];
        if (e == 0xFFFF) {
          d[0] = 0xFFFF;
          continue;
        }
        const uint16_t* c = P.my_cnt[q].data() + (size_t)(P.ent_off[q][b] + e) * 6;
        for (int k = 0; k < 6; k++) d[k] = c[k];
      }
  }
  return "";
}

uint32_t default_run_length(const Plan& P) {
  // long runs amortize the tables; short ones keep a small tree's
  // workgroups spread over the chip (about 2048 runs at least: the turn's
  // 10,526 terminals in runs of 5, terminal pass 0.108 -> 0.098 ms per
  // iteration against runs of 2; runs of 10 gain nothing)
  return std::min<uint32_t>(kBatchMax, std::max<uint32_t>(1, (uint32_t)(P.trec.size() / 2048)));
}

void build_runs(Plan& P, uint32_t bmax) {
  bmax = std::min(std::max(bmax, 1u), kBatchMax);
  const uint32_t C = P.num_chunks;
  P.tbatch.clear();
  P.tb_coff.assign(C + 1, 0);
  auto runs = [&](uint32_t a, uint32_t b) {
    uint32_t t = a;
    while (t < b) {
      TermBatch r{t, 1, P.trec[t].board, P.trec[t].spot};
      while (t + r.cnt < b && r.cnt < bmax && P.trec[t + r.cnt].board == r.board && P.trec[t + r.cnt].spot == r.spot) r.cnt++;
      P.tbatch.push_back(r);
      t += r.cnt;
    }
  };
  runs(0, C ? P.term_coff[0] : (uint32_t)P.trec.size());
  for (uint32_t c = 0; c < C; c++) {
    P.tb_coff[c] = (uint32_t)P.tbatch.size();
    runs(P.term_coff[c], P.term_coff[c + 1]);
  }
  P.tb_coff[C] = (uint32_t)P.tbatch.size();
}

std::string merge_plans(std::vector<Plan>&& parts, Plan& M) {
  const uint32_t K = (uint32_t)parts.size();
  if (K == 0) return fail("empty batch", 0);
  if (K > 0xFFFF) return fail("too many spots in a batch", K);
  M = Plan();
  // offsets of every spot's ids in the merged numbering
  std::vector<uint64_t> node_off(K), vslot_off(K), store_off(K), term_off(K), slot_off[2];
  slot_off[0].resize(K);
  slot_off[1].resize(K);
  uint64_t nn = 0, nv = 0, ns = 0, nt = 0, sl[2] = {0, 0}, nb = 0, nps = 0;
  std::vector<uint64_t> ps_off(K);
  bool q16 = true;
  uint32_t D = 0;
  for (uint32_t k = 0; k < K; k++) {
    const Plan& T = parts[k];
    node_off[k] = nn;
    vslot_off[k] = nv;
    store_off[k] = ns;
    term_off[k] = nt;
    ps_off[k] = nps;
    nps += T.num_pslots;
    q16 = q16 && T.num_pslots > 0;
    for (int p = 0; p < 2; p++) {
      slot_off[p][k] = sl[p];
      sl[p] += T.num_slots[p];
    }
    nn += T.num_nodes;
    nv += T.num_vslots;
    ns += T.storage_len;
    nt += T.trec.size();
    nb += T.num_boards;
    D = std::max(D, T.depths());
    M.n[0] = std::max(M.n[0], T.n[0]);
    M.n[1] = std::max(M.n[1], T.n[1]);
    M.max_members = std::max(M.max_members, T.max_members);
    M.max_entries = std::max(M.max_entries, T.max_entries);
  }
  if (nn >= 0xFFFFFFFFull || nv >= 0xFFFFFFFFull || sl[0] >= 0xFFFFFFFFull || sl[1] >= 0xFFFFFFFFull || nt >= 0xFFFFFFFFull)
    return fail("batch too large for 32-bit ids", nn);
  M.num_nodes = nn;
  M.num_vslots = (uint32_t)nv;
  M.storage_len = ns;
  M.num_pslots = q16 && nps < 0xFFFFFFFFull ? (uint32_t)nps : 0;
  M.num_slots[0] = (uint32_t)sl[0];
  M.num_slots[1] = (uint32_t)sl[1];
  M.num_boards = (uint32_t)nb;
  M.kind.reserve(nn);
  M.na.reserve(nn);
  M.card.reserve(nn);
  for (uint32_t k = 0; k < K; k++) {
    const Plan& T = parts[k];
    M.kind.insert(M.kind.end(), T.kind.begin(), T.kind.end());
    M.na.insert(M.na.end(), T.na.begin(), T.na.end());
    M.card.insert(M.card.end(), T.card.begin(), T.card.end());
  }
  auto shift = [&](MemRec r, uint32_t k) {
    r.node += (uint32_t)node_off[k];
    if (r.vslot != kNone) r.vslot += (uint32_t)vslot_off[k];
    for (int p = 0; p < 2; p++) {
      if (r.slot[p] != kNone) r.slot[p] += (uint32_t)slot_off[p][k];
      if (r.cslot[p] != kNone) r.cslot[p] += (uint32_t)slot_off[p][k];
    }
    if (r.kind == kChance) r.offset += node_off[k];  // first child's node id
    else if (is_player(r.kind)) {
      r.offset += store_off[k];
      r.pslot += (uint32_t)ps_off[k];
    }
    r.spot = (uint16_t)k;
    return r;
  };
  // segments by depth, spot by spot within a depth
  M.depth_off.assign(D + 1, 0);
  std::vector<uint32_t> pslot[2];
  for (uint32_t d = 0; d < D; d++) {
    M.depth_off[d] = (uint32_t)M.seg_root.size();
    for (uint32_t k = 0; k < K; k++) {
      const Plan& T = parts[k];
      if (d >= T.depths()) continue;
      const size_t S = T.seg_root.size();
      for (uint32_t g = T.depth_off[d]; g < T.depth_off[d + 1]; g++) {
        M.seg_root.push_back(T.seg_root[g] + (uint32_t)node_off[k]);
        M.seg_mem_off.push_back((uint32_t)M.mrec.size());
        M.seg_mem_cnt.push_back(T.seg_mem_cnt[g]);
        for (uint32_t m = 0; m < T.seg_mem_cnt[g]; m++) M.mrec.push_back(shift(T.mrec[T.seg_mem_off[g] + m], k));
        for (int p = 0; p < 2; p++) {
          uint32_t x = T.seg_pslot[(size_t)p * S + g];
          pslot[p].push_back(x == kNone ? kNone : x + (uint32_t)slot_off[p][k]);
        }
        M.seg_factor.push_back(T.seg_factor[g]);
      }
    }
  }
  M.depth_off[D] = (uint32_t)M.seg_root.size();
  M.seg_pslot = pslot[0];
  M.seg_pslot.insert(M.seg_pslot.end(), pslot[1].begin(), pslot[1].end());
  // chance records by depth
  M.chance_off.assign(D + 1, 0);
  M.ciso_off.assign(1, 0);
  M.ciso.clear();
  for (uint32_t d = 0; d < D; d++) {
    M.chance_off[d] = (uint32_t)M.crec.size();
    for (uint32_t k = 0; k < K; k++) {
      const Plan& T = parts[k];
      if (d >= T.depths()) continue;
      for (uint32_t i = T.chance_off[d]; i < T.chance_off[d + 1]; i++) {
        M.crec.push_back(shift(T.crec[i], k));
        M.chance_child_vslot.push_back(T.chance_child_vslot[i] + (uint32_t)vslot_off[k]);
        M.ciso.insert(M.ciso.end(), T.ciso.begin() + T.ciso_off[i], T.ciso.begin() + T.ciso_off[i + 1]);
        M.ciso_off.push_back((uint32_t)M.ciso.size());
      }
    }
  }
  M.chance_off[D] = (uint32_t)M.crec.size();
  // terminals, spot by spot
  for (uint32_t k = 0; k < K; k++) {
    const Plan& T = parts[k];
    for (size_t t = 0; t < T.trec.size(); t++) {
      TermRec r = T.trec[t];
      r.node += (uint32_t)node_off[k];
      r.slot[0] += (uint32_t)slot_off[0][k];
      r.slot[1] += (uint32_t)slot_off[1][k];
      r.vslot += (uint32_t)vslot_off[k];
      r.spot = (uint16_t)k;
      M.trec.push_back(r);
      M.terminals.push_back(T.terminals[t] + (uint32_t)node_off[k]);
    }
  }
  // one resident group: no chunks
  M.num_chunks = 0;
  M.chunk_of.assign(nn, kNone);
  M.seg_coff.assign(D, 0);
  M.ch_coff.assign(D, 0);
  M.term_coff.assign(1, (uint32_t)M.trec.size());
  for (int p = 0; p < 2; p++) {
    M.slot_s0[p] = M.num_slots[p];
    M.slot_coff[p].assign(1, M.num_slots[p]);
  }
  M.val_v0 = M.num_vslots;
  M.val_coff.assign(1, M.num_vslots);
  build_runs(M, default_run_length(M));
  // spot roots
  M.num_spots = K;
  M.spot_root_node.resize(K);
  M.spot_root_vslot.resize(K);
  M.spot_root_slot[0].resize(K);
  M.spot_root_slot[1].resize(K);
  for (uint32_t k = 0; k < K; k++) {
    M.spot_root_node[k] = (uint32_t)node_off[k];
    M.spot_root_vslot[k] = parts[k].spot_root_vslot[0] + (uint32_t)vslot_off[k];
    for (int p = 0; p < 2; p++) M.spot_root_slot[p][k] = parts[k].spot_root_slot[p][0] + (uint32_t)slot_off[p][k];
  }
  M.half_pot = parts[0].half_pot;
  M.norm = parts[0].norm;
  M.batch_river = parts[0].batch_river;
  M.parts = std::move(parts);
  return "";
}

ScratchRows partition_chunks(const Plan& P, size_t es, uint32_t rowlen, uint64_t budget, std::vector<ChunkGroup>& G) {
  const uint32_t C = P.num_chunks;
  auto rows = [&](uint32_t c0, uint32_t c1, ScratchRows& r) {
    r.reach_bytes = 0;
    r.alive_rows = 0;
    for (int p = 0; p < 2; p++) {
      const uint64_t n = (uint64_t)P.slot_s0[p] + (C ? P.slot_coff[p][c1] - P.slot_coff[p][c0] : 0);
      r.slot_rows[p] = n;
      r.reach_bytes = std::max<uint64_t>(r.reach_bytes, n * rowlen * es);
      r.alive_rows = std::max<uint64_t>(r.alive_rows, n);
    }
    r.val_rows = (uint64_t)P.val_v0 + (C ? P.val_coff[c1] - P.val_coff[c0] : 0);
    r.val_bytes = r.val_rows * rowlen * es;
    return r.reach_bytes + r.val_bytes + r.alive_rows * 4;
  };
  ScratchRows best, cur;
  auto take = [&](const ScratchRows& r) {
    best.reach_bytes = std::max(best.reach_bytes, r.reach_bytes);
    best.val_bytes = std::max(best.val_bytes, r.val_bytes);
    best.alive_rows = std::max(best.alive_rows, r.alive_rows);
    for (int p = 0; p < 2; p++) best.slot_rows[p] = std::max(best.slot_rows[p], r.slot_rows[p]);
    best.val_rows = std::max(best.val_rows, r.val_rows);
  };
  G.clear();
  if (C == 0) {
    rows(0, 0, cur);
    take(cur);
    G.push_back({0, 0});
    return best;
  }
  // the allocation is the largest reach, value and alive rows over all
  // groups, so a group grows only while that combination fits
  auto total = [&](const ScratchRows& r) {
    return std::max(best.reach_bytes, r.reach_bytes) + std::max(best.val_bytes, r.val_bytes) +
           std::max(best.alive_rows, r.alive_rows) * 4;
  };
  uint32_t c0 = 0;
  while (c0 < C) {
    uint32_t c1 = c0 + 1;
    rows(c0, c1, cur);
    ScratchRows nx;
    while (c1 < C) {
      rows(c0, c1 + 1, nx);
      if (total(nx) > budget) break;
      cur = nx;
      c1++;
    }
    take(cur);
    G.push_back({c0, c1});
    c0 = c1;
  }
  return best;
}

std::string check_groups(const Plan& P, const std::vector<ChunkGroup>& G, const ScratchRows& R) {
  const uint32_t C = P.num_chunks, D = P.depths(), N = (uint32_t)P.num_nodes;
  char b[200];
  for (int p = 0; p < 2; p++) {
    const uint32_t s0 = P.slot_s0[p], v0 = P.val_v0;
    auto check = [&](uint32_t s, uint32_t n, bool val, uint32_t c0, uint32_t c1, const char* what) -> bool {
      if (s == kNone) return true;
      const uint32_t r0 = val ? v0 : s0;
      const uint32_t lo = C ? (val ? P.val_coff[c0] : P.slot_coff[p][c0]) : r0;
      const uint32_t hi = C ? (val ? P.val_coff[c1] : P.slot_coff[p][c1]) : r0;
      const uint64_t rows = val ? R.val_rows : R.slot_rows[p];
      bool ok;
      if ((uint64_t)s + n <= r0) ok = true;  // resident
      else ok = s >= r0 && s >= lo && (uint64_t)s + n <= hi && (uint64_t)(s - lo) + r0 + n <= rows;
      if (!ok) std::snprintf(b, sizeof b, "group scratch: %s %u (+%u) of traverser %d outside chunks [%u, %u)", what, s, n, p, c0, c1);
      return ok;
    };
    auto seg_ok = [&](uint32_t g, uint32_t c0, uint32_t c1, bool top) -> bool {
      if (!check(P.seg_pslot[(size_t)p * P.seg_root.size() + g], 1, false, c0, c1, "parent slot")) return false;
      for (uint32_t m = 0; m < P.seg_mem_cnt[g]; m++) {
        const MemRec& r = P.mrec[P.seg_mem_off[g] + m];
        if (top && P.chunk_of[r.node] != kNone) {
          std::snprintf(b, sizeof b, "group scratch: node %u of the top segment is in a chunk", r.node);
          return false;
        }
        if (!check(r.slot[p], 1, false, c0, c1, "slot")) return false;
        if (r.cslot[p] != kNone && !check(r.cslot[p], r.na, false, c0, c1, "child slots")) return false;
        if (r.vslot != kNone && !check(r.vslot, 1, true, c0, c1, "value row")) return false;
      }
      return true;
    };
    auto term_ok = [&](uint32_t t, uint32_t c0, uint32_t c1) {
      const TermRec& r = P.trec[t];
      return check(r.slot[p], 1, false, c0, c1, "terminal slot") && check(r.vslot, 1, true, c0, c1, "terminal row");
    };
    auto chance_ok = [&](uint32_t i, uint32_t c0, uint32_t c1) {
      const MemRec& r = P.crec[i];
      return check(r.slot[p], 1, false, c0, c1, "chance slot") && check(r.vslot, 1, true, c0, c1, "chance row") &&
             check(P.chance_child_vslot[i], r.na, true, c0, c1, "chance child rows");
    };
    // the top segment and its chance records use resident rows only
    for (uint32_t g = P.depth_off[0]; g < P.depth_off[1]; g++)
      if (!seg_ok(g, 0, 0, true)) return std::string(b);
    for (uint32_t i = P.chance_off[0]; i < P.chance_off[1]; i++)
      if (!chance_ok(i, 0, 0)) return std::string(b);
    if (C == 0) {  // every row resident: every segment and chance record at every depth
      for (uint32_t g = P.depth_off[1]; g < P.depth_off[D]; g++)
        if (!seg_ok(g, 0, 0, false)) return std::string(b);
      for (uint32_t i = P.chance_off[1]; i < P.chance_off[D]; i++)
        if (!chance_ok(i, 0, 0)) return std::string(b);
      for (uint32_t t = 0; t < (uint32_t)P.trec.size(); t++)
        if (!term_ok(t, 0, 0)) return std::string(b);
      uint32_t t = 0;
      for (const TermBatch& r : P.tbatch) {
        if (r.t0 != t || r.cnt == 0 || r.cnt > kBatchMax) return fail("terminal runs do not tile the tree", t);
        t += r.cnt;
      }
      if (t != (uint32_t)P.trec.size()) return fail("terminal runs do not tile the tree", t);
      continue;
    }
    std::vector<uint8_t> seen_t(P.trec.size(), 0);
    for (const ChunkGroup& g : G) {
      for (uint32_t d = 1; d < D; d++) {
        const uint32_t* so = P.seg_coff.data() + (size_t)d * (C + 1);
        const uint32_t* co = P.ch_coff.data() + (size_t)d * (C + 1);
        for (uint32_t s = so[g.c0]; s < so[g.c1]; s++)
          if (!seg_ok(s, g.c0, g.c1, false)) return std::string(b);
        for (uint32_t i = co[g.c0]; i < co[g.c1]; i++)
          if (!chance_ok(i, g.c0, g.c1)) return std::string(b);
      }
      const uint32_t t0 = g.c0 == 0 ? 0u : P.term_coff[g.c0];
      for (uint32_t t = t0; t < P.term_coff[g.c1]; t++) {
        if (!term_ok(t, g.c0, g.c1)) return std::string(b);
        seen_t[t] = 1;
      }
    }
    for (size_t t = 0; t < seen_t.size(); t++)
      if (!seen_t[t]) return fail("terminal outside every group", t);
    // the terminal runs of a group cover exactly its terminals
    for (const ChunkGroup& g : G) {
      const uint32_t b0 = g.c0 == 0 ? 0u : P.tb_coff[g.c0], b1 = P.tb_coff[g.c1];
      uint32_t t = g.c0 == 0 ? 0u : P.term_coff[g.c0];
      for (uint32_t k = b0; k < b1; k++) {
        const TermBatch& r = P.tbatch[k];
        if (r.t0 != t || r.cnt == 0 || r.cnt > kBatchMax) return fail("terminal runs do not tile the group", k);
        for (uint32_t j = r.t0; j < r.t0 + r.cnt; j++)
          if (P.trec[j].board != r.board) return fail("terminal run with two boards", j);
        t += r.cnt;
      }
      if (t != P.term_coff[g.c1]) return fail("terminal runs do not tile the group", b1);
    }
    (void)N;
  }
  return "";
}

}  // namespace pts
