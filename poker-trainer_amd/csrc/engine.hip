// HIP kernels for gfx950 and the engine that drives them.
//
// Every arithmetic step repeats the reference CPU solve's operation order
// (crates/solver/src/cfr.rs, eval.rs, river.rs) with IEEE f32/f64 operations
// and no contraction (built with -ffp-contract=off), and every f64 sum that
// the reference accumulates sequentially is accumulated sequentially here
// too (one lane per chain), so the engine's regrets, strategy sums and values
// are bit-identical to the CPU solve's.
//
// Storage layout (shared with the caller): node block at `offset`,
// `action * hands + hand` (crates/solver/src/storage.rs:87-106). Threads run
// over hands, so every storage access is a coalesced row access.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>

#include "engine.hpp"
#include "jit.hpp"
#include "net.hpp"

namespace pts {

static thread_local std::string g_err;
static thread_local ErrorDetail g_detail;
void set_error(const std::string& m) {
  g_err = m;
  g_detail = ErrorDetail{};
}
const std::string& last_error() { return g_err; }
void set_error_detail(const ErrorDetail& d) { g_detail = d; }
const ErrorDetail& last_error_detail() { return g_detail; }

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) {                                                                        \
      set_error(std::string("HIP error in ") + #x + ": " + hipGetErrorString(e_));                 \
      set_error_detail(ErrorDetail{(int32_t)e_, 0, 0});                                            \
      return e_ == hipErrorOutOfMemory ? kOutOfMemory : kDriver;                                   \
    }                                                                                              \
  } while (0)

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint16_t NONE16 = 0xFFFF;
constexpr int MAXA = 8;       // actions per decision node (checked at create)
constexpr int TERM_T = 64;    // terminal workgroup: one wave
constexpr int UP_T = 64;      // up workgroup: one wave, one hand per lane
constexpr int DOWN_T = 512;   // down workgroup
constexpr int MAX_K = 21;     // ceil(1326 / 64): hands per lane in the terminal kernel
constexpr int DG_A = 16;      // actions per down-pass group
// actions per down-pass group with 16-bit storage, whose down pass stages a
// group's new sums in LDS (one row per action): 8 rows let three 512-thread
// workgroups share a CU where 16 let two (the flop's 16-bit down pass 131.6 ->
// 101.2 ms); a node has at most MAXA = 8 actions, so one node always fits
constexpr int DG_A16 = 8;
static_assert(DG_A16 >= MAXA && DG_A16 <= DG_A, "one node's actions fit a 16-bit group's LDS rows");
constexpr int DG_N = 8;       // nodes per down-pass group

// One spot's hands and strength tables on the device (a batch holds many
// spots; every record names its spot).
struct SpotTab {
  uint32_t n[2];
  uint32_t root_slot[2];  // the root's reach slot, per traverser
  uint32_t root_vslot, pad;
  double half_pot, norm;
  const float* weights[2];
  const uint8_t* cards[2];
  const uint32_t* same[2];
  const uint32_t* bl_off[2];
  const uint16_t* bl_idx[2];
  const uint32_t *ent_off[2], *ent_cnt[2], *cl_base[2];
  const uint16_t *ent_idx[2], *cl_off[2], *cl_idx[2];
  const uint4* hand_cnt[2];
  const uint32_t* lookup[2];  // combo -> hand, or NONE
  const uint16_t* perm[2];    // 6 suit-pair hand permutations of n each (null without isomorphism)
};

// Device view of the plan (all pointers into one HBM blob).
struct DPlan {
  uint32_t num_nodes, num_segs, num_spots;
  uint32_t max_entries, max_members;
  uint32_t max_dec;  // the most decision members in a segment
  const int8_t* card;
  const uint32_t *seg_mem_off, *seg_mem_cnt, *seg_pslot;
  const float* seg_factor;
  const MemRec* mrec;
  const TermRec* trec;
  const TermBatch* tbatch;
  const MemRec* crec;
  const uint32_t* chance_child_vslot;
  const uint32_t* ciso_off;  // per chance record: its cards merged by isomorphism, ciso[ciso_off[i], ciso_off[i + 1])
  const IsoRef* ciso;
  const SpotTab* spots;
  const uint32_t* active;  // per spot: 0 freezes the spot (its root is dead)
  // per segment: {first member, member count, root node, hands of player
  // 0 | hands of player 1 << 16}
  const uint4* seg_head;
  // per member: its alive flag for the current traverser (written by the
  // down pass, read by the up pass)
  uint8_t* malive;
  // segments of more than REG_L members, of REG_M + 1 to REG_L, of REG_S + 1
  // to REG_M and of REG_T + 1 to REG_S members, and the segments with a walk
  // of up to REG_T and up to REG_S members (ascending): device copies, and
  // the host copies launches are cut from. A chance child whose one member
  // is a terminal or a chance node has no walk: that member writes the
  // segment root's row itself, and a chance node reads a dead child as zero.
  const uint32_t *big, *m80, *mid, *m48, *w32, *w48;
  const uint32_t *big_host, *m80_host, *mid_host, *m48_host, *w32_host, *w48_host;
  uint32_t num_big, num_m80, num_mid, num_m48, num_w32, num_w48;
  // per traverser p: each segment's opponent decision nodes in groups the
  // down pass runs together (one level of the segment, at most DG_A actions
  // in all): entries dg_ent[p][dg_off[p][seg] .. dg_off[p][seg + 1]), each
  // the member's local index | its group's number << 16
  const uint32_t* dg_off[2];
  const uint32_t* dg_ent[2];
};

// The nodes a traversal covers: {root} and [lo, hi) (the whole tree, or the
// subtree below a queried decision node).
// A whole-tree traversal (init = 1) starts every spot's root reach row from
// the opponent's range in its down pass and copies every spot's root value
// row into res[(spot * 2 + p) * rowlen] in its up pass.
struct Span {
  uint32_t root, lo, hi, seg;  // seg: segment holding `root`
  uint32_t init;
  void* res;
  // a query's node's children [kid0, kid1): the query traverses each child on
  // its own (lib.rs:854-863), so a terminal child is evaluated by itself, not
  // as part of its river node's batch (cfr.rs:277-281)
  uint32_t kid0 = 0, kid1 = 0;
  __device__ bool has(uint32_t id) const { return id == root || (id >= lo && id < hi); }
  __device__ bool kid(uint32_t id) const { return id >= kid0 && id < kid1; }
};

// Reach slots and value rows of one chunk group: resident ones (below s0 /
// v0) are used as numbered, the group's own are shifted onto the shared
// scratch rows that follow the resident ones.
struct Bias {
  uint32_t s0, sb, v0, vb;
  __device__ __forceinline__ uint32_t s(uint32_t x) const { return x < s0 ? x : x - sb; }
  __device__ __forceinline__ uint32_t v(uint32_t x) const { return x < v0 ? x : x - vb; }
};

// Regret and strategy-sum storage (crates/solver/src/storage.rs:87-106):
// f32 cells, or 16-bit cells (i16 regrets, u16 sums) with one f32 scale per
// player node, a stored q standing for q * (scale / MAX).
struct Store {
  float* r;      // f32 regrets (null with 16-bit storage)
  float* s;      // f32 strategy sums
  int16_t* rq;   // 16-bit regrets
  uint16_t* sq;  // 16-bit strategy sums
  float* rsc;    // per player node: the regrets' scale
  float* ssc;    // per player node: the strategy sums' scale
  int32_t* rnx;  // per player node: the regrets' new scale (float bits) while a CFR traversal runs, else -1
};
constexpr float kRegMax = 32767.0f, kSumMax = 65535.0f;  // i16::MAX, u16::MAX as f32
// Rust's saturating `f32 as i16` / `as u16`: truncation toward zero, NaN to 0
__device__ __forceinline__ int16_t sat_i16(float y) {
  if (!(y == y)) return 0;
  return (int16_t)(int32_t)fminf(fmaxf(y, -32768.0f), 32767.0f);
}
__device__ __forceinline__ uint16_t sat_u16(float y) {
  if (!(y == y)) return 0;
  return (uint16_t)(uint32_t)fminf(fmaxf(y, 0.0f), 65535.0f);
}
// storage.rs:274-281: y = v * (MAX / max), rounded half away from zero
__device__ __forceinline__ int16_t enc_regret(float v, float inv) {
  const float y = v * inv;
  return sat_i16(y + copysignf(0.5f, y));
}
// storage.rs:312-317: max(v, 0) * (MAX / max) + 0.5, truncated
__device__ __forceinline__ uint16_t enc_sum(float v, float inv) { return sat_u16(fmaxf(v, 0.0f) * inv + 0.5f); }
// The largest of the wave's 64 values, all >= 0 (0 is the identity), in
// every lane: inclusive max-scans within each row of 16 lanes (DPP
// row_shr 1, 2, 4, 8), row 0's and row 2's last lanes into rows 1 and 3
// (row_bcast:15), then row 1's last lane into rows 2 and 3 (row_bcast:31):
// lane 63 ends with the maximum. Lanes a DPP move leaves without a source
// take 0.
__device__ __forceinline__ float dpp_max(float x, int src) { return fmaxf(x, __int_as_float(src)); }
__device__ __forceinline__ float wave_max(float x) {
  x = dpp_max(x, __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xF, 0xF, false));  // row_shr:1
  x = dpp_max(x, __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x112, 0xF, 0xF, false));  // row_shr:2
  x = dpp_max(x, __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x114, 0xF, 0xF, false));  // row_shr:4
  x = dpp_max(x, __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x118, 0xF, 0xF, false));  // row_shr:8
  x = dpp_max(x, __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x142, 0xA, 0xF, false));  // row_bcast:15
  x = dpp_max(x, __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x143, 0xC, 0xF, false));  // row_bcast:31
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Per-player fields of a record held in registers: selects, not a runtime
// index (which would put the record in scratch memory).
__device__ __forceinline__ uint32_t pick(const uint32_t (&a)[2], int p) { return p ? a[1] : a[0]; }
__device__ __forceinline__ int32_t pick(const int32_t (&a)[2], int p) { return p ? a[1] : a[0]; }

template <typename V> __device__ __forceinline__ V pos0(V x) { return x > V(0) ? x : V(0); }
template <typename V> __device__ __forceinline__ V recip(V d) {  // cfr.rs:113-121
  V i = V(1) / d;
  return (d > V(0) && isfinite(i)) ? i : V(0);
}
__device__ __forceinline__ bool holds(const uint8_t* __restrict__ cards, uint32_t h, int c) {
  return cards[2 * h] == c || cards[2 * h + 1] == c;
}

__device__ __forceinline__ uint32_t block_or(uint32_t bits, uint32_t* red) {
  __syncthreads();
  if (threadIdx.x == 0) *red = 0;
  __syncthreads();
  if (bits) atomicOr(red, bits);
  __syncthreads();
  return *red;
}

// Copies a segment's member records into LDS (all threads of the block).
__device__ __forceinline__ void stage_records(const MemRec* __restrict__ g, uint32_t cnt, MemRec* lds) {
  const uint4* src = (const uint4*)g;
  uint4* dst = (uint4*)lds;
  for (uint32_t i = threadIdx.x; i < cnt * 3; i += blockDim.x) dst[i] = src[i];
}

// ------------------------------------------------------------------ begin
// One reach row for a query: the caller's opponent reach at `slot`.
template <typename V>
__global__ void begin_row_kernel(uint32_t nq, const float* __restrict__ w, V* __restrict__ reach, uint32_t slot,
                                 uint32_t* __restrict__ alive, uint32_t is_alive, uint32_t rowlen) {
  const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < nq) reach[(size_t)slot * rowlen + h] = (V)w[h];
  if (h == 0) alive[slot] = is_alive;
}

// ------------------------------------------------------------------ down
// One workgroup per segment: the opponent's reach pushed down the segment's
// opponent decision nodes, a level at a time. CFR mode (V = float): regret matching and
// the strategy-sum update (cfr.rs:156-207,389-419; river.rs:116-160).
// Otherwise the normalized average strategy (cfr.rs:124-154).
template <typename V, int MODE, bool Q>
__global__ void __launch_bounds__(1024) down_kernel(DPlan P, int p, uint32_t seg0, Span span, Bias B, V* __restrict__ reach,
                                                      uint32_t* __restrict__ alive, Store sto, const IterParams* __restrict__ prm,
                                                      uint32_t rowlen) {
  extern __shared__ uint4 smem_down[];
  __shared__ uint32_t red;
  const uint32_t seg = seg0 + blockIdx.x;
  const int o = p ^ 1;
  const uint32_t moff = P.seg_mem_off[seg], mcnt = P.seg_mem_cnt[seg];
  const uint32_t root = P.mrec[moff].node;
  const SpotTab& S = P.spots[P.mrec[moff].spot];
  const uint32_t nq = S.n[o], tid = threadIdx.x;
  if (!span.has(root) && seg != span.seg) return;
  MemRec* rec = (MemRec*)smem_down;
  uint8_t* m_alive = (uint8_t*)(rec + P.max_members);
  // 16-bit storage: a group's new sums (its nodes' rows, action-major, nq
  // per row) between their maxima and their encoding (after the staged
  // group entries)
  float* ybuf = (float*)(m_alive + ((P.max_members + 15) & ~15u) + 4 * (size_t)P.max_members);
  stage_records(P.mrec + moff, mcnt, rec);
  __syncthreads();
  const uint32_t lane = tid & 63;
  const uint32_t sp_raw = P.seg_pslot[(size_t)p * P.num_segs + seg];
  if (span.init && sp_raw == NONE) {
    // a spot's root: its reach row is the opponent's range weights, alive
    // unless the spot is frozen
    const uint32_t slot = S.root_slot[p];
    for (uint32_t h = tid; h < nq; h += blockDim.x) reach[(size_t)slot * rowlen + h] = (V)S.weights[o][h];
    if (tid == 0) {
      const uint32_t on = P.active[P.mrec[moff].spot];
      alive[slot] = on;
      m_alive[rec[0].src[p]] = (uint8_t)(on != 0);
    }
  } else if (seg == span.seg || sp_raw == NONE) {
    if (tid == 0) {
      // the span root's (or a spot root's) row was written by a begin kernel:
      // a spot's root segment starts at member 0; the span's segment at the
      // member holding span.root (the host checks that it is one)
      uint32_t rl = 0;
      if (seg == span.seg && span.root != root)
        for (uint32_t m = 1; m < mcnt; m++)
          if (rec[m].node == span.root) {
            rl = m;
            break;
          }
      m_alive[rec[rl].src[p]] = (uint8_t)(alive[B.s(rec[rl].slot[p])] != 0);
    }
  } else {
    // Chance child: the parent's reach times the chance factor, opponent
    // hands holding the dealt card zeroed (cfr.rs:449-456).
    const uint32_t sp = B.s(sp_raw), sm = B.s(rec[0].slot[p]);
    uint32_t any = 0;
    if (alive[sp]) {
      const V f = (V)P.seg_factor[seg];
      const int c = rec[0].card;
      const V* pr = reach + (size_t)sp * rowlen;
      V* mr = reach + (size_t)sm * rowlen;
      for (uint32_t h = tid; h < nq; h += blockDim.x) {
        V v = pr[h] * f;
        if (holds(S.cards[o], h, c)) v = V(0);
        mr[h] = v;
        any |= v != V(0);
      }
    }
    any = block_or(any, &red);
    if (tid == 0) {
      alive[sm] = any;
      m_alive[0] = (uint8_t)any;
    }
  }
  __syncthreads();
  float g = 0.0f;
  bool weighted = false;
  if (MODE == kCfr) {
    g = prm->strat;
    weighted = prm->reach_weighted != 0;
  }
  // The opponent's decision nodes in groups of one level (their reach rows
  // were written by earlier groups): the waves share the group's
  // (node, 64-hand tile) units, two units' loads in flight per wave, then
  // one barrier settles every node's children alive bits (a ballot per
  // action per unit, an LDS OR per node) and, with 16-bit storage, every
  // node's largest new sum.
  __shared__ uint32_t gor[3][DG_N];  // per group node: its children's alive bits, buffers used in turn
  __shared__ int32_t gmx[3][DG_N];   // 16-bit storage: per group node: the largest |new sum| (float bits)
  uint32_t* ents = (uint32_t*)(m_alive + ((P.max_members + 15) & ~15u));  // the segment's groups, staged
  const uint32_t e0 = P.dg_off[p][seg], ne = P.dg_off[p][seg + 1] - e0;
  for (uint32_t k = tid; k < ne; k += blockDim.x) ents[k] = P.dg_ent[p][e0 + k];
  if (tid < 3 * DG_N) {
    (&gor[0][0])[tid] = 0;
    (&gmx[0][0])[tid] = 0;
  }
  __syncthreads();
  const uint32_t wave = tid >> 6, nw = blockDim.x >> 6;
  const uint32_t tiles = (nq + 63) >> 6;
  uint32_t gi = 0;
  for (uint32_t e = 0; e < ne;) {
    const uint32_t gid = ents[e] >> 16;
    uint32_t gn = 1;
    while (e + gn < ne && (ents[e + gn] >> 16) == gid) gn++;
    const uint32_t units = gn * tiles;
    for (uint32_t u0 = wave; u0 < units; u0 += 2 * nw) {
      constexpr int UB = 2;
      // the units' nodes, hands and inputs, every load issued before any store
      uint32_t jj[UB], hx[UB], na[UB];
      bool on[UB], act[UB];
      uint64_t off[UB];
      const MemRec* rp[UB];
      V rr[UB];
      float rg[UB][MAXA], xs[UB][MAXA];
      V sv[UB][MAXA];
#pragma unroll
      for (int k = 0; k < UB; k++) {
        const uint32_t u = u0 + k * nw;
        const uint32_t j = u < units ? u / tiles : 0;
        const MemRec& r = rec[ents[e + j] & 0xFFFF];
        rp[k] = &r;
        jj[k] = j;
        on[k] = u < units && span.has(r.node) && m_alive[r.src[p]];  // dead: no rows, no updates (cfr.rs:271-274)
        const uint32_t h = (u - j * tiles) * 64 + lane;
        act[k] = on[k] && h < nq;
        hx[k] = act[k] ? h : 0;
        na[k] = on[k] ? r.na : 0;
        off[k] = r.offset;
        rr[k] = on[k] ? reach[(size_t)B.s(r.slot[p]) * rowlen + hx[k]] : V(0);
        const float sstep = Q && on[k] ? sto.ssc[r.pslot] / kSumMax : 0.0f;
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na[k]) {
            const size_t i = off[k] + (size_t)a * nq + hx[k];
            if (MODE == kCfr) rg[k][a] = Q ? (float)sto.rq[i] : sto.r[i];
            xs[k][a] = Q ? (float)sto.sq[i] * sstep : sto.s[i];
          }
      }
      // f32 DCFR rounds compute every output before the first global store:
      // a store ahead of a read of a loaded value would make that read wait
      // for the store's completion (the other modes store as they go, which
      // needs fewer registers)
      constexpr bool DEFER = std::is_same<V, float>::value && MODE == kCfr && !Q;
      float yv[UB][MAXA];
      V cv[UB][MAXA];
      uint32_t wbk[UB];
      float lmx[UB];
#pragma unroll
      for (int k = 0; k < UB; k++) {
        wbk[k] = 0;
        lmx[k] = 0.0f;
        if (!on[k]) continue;
        const uint32_t n = na[k], h = hx[k];
        V sa[MAXA];
        float lmax = 0.0f;
        if (MODE == kCfr) {
          // cfr.rs:389-414: regret matching (16-bit regrets undecoded, the
          // scale cancels: storage.rs:221-234), then the sums x * g + s (* reach)
          float d = pos0(rg[k][0]);
#pragma unroll
          for (int a = 1; a < MAXA; a++)
            if (a < (int)n) d += pos0(rg[k][a]);
          const float inv = recip(d), uni = 1.0f / (float)n;
          uint32_t yo = 0;  // this node's rows in the group's staging rows
          if (Q)
            for (uint32_t q = 0; q < jj[k]; q++) yo += rec[ents[e + q] & 0xFFFF].na;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)n) {
              const float s = inv == 0.0f ? uni : pos0(rg[k][a]) * inv;
              const float x = xs[k][a];
              const float y = weighted ? x * g + s * (float)rr[k] : x * g + s;
              if (Q) {
                if (act[k]) {
                  lmax = fmaxf(lmax, fabsf(y));
                  ybuf[(yo + a) * nq + h] = y;
                }
              } else if (DEFER) {
                yv[k][a] = y;
              } else if (act[k]) {
                sto.s[off[k] + (size_t)a * nq + h] = y;
              }
              sa[a] = (V)s;
            }
        } else {
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)n) sv[k][a] = (V)xs[k][a];
          V d = sv[k][0];
#pragma unroll
          for (int a = 1; a < MAXA; a++)
            if (a < (int)n) d += sv[k][a];
          const V inv = recip(d), uni = V(1) / (V)n;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)n) sa[a] = inv == V(0) ? uni : sv[k][a] * inv;
        }
        uint32_t wb = 0;
        V* crow = reach + (size_t)B.s(rp[k]->cslot[p]) * rowlen;
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)n) {
            const V v = sa[a] * rr[k];
            if (DEFER) cv[k][a] = v;
            else if (act[k]) crow[(size_t)a * rowlen + h] = v;
            if (__ballot(act[k] && v != V(0))) wb |= 1u << a;
          }
        wbk[k] = wb;
        lmx[k] = lmax;
      }
#pragma unroll
      for (int k = 0; k < UB; k++) {
        if (!on[k]) continue;
        const MemRec& r = *rp[k];
        const uint32_t n = na[k], h = hx[k];
        if (DEFER && act[k]) {
          V* crow = reach + (size_t)B.s(r.cslot[p]) * rowlen;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)n) {
              sto.s[off[k] + (size_t)a * nq + h] = yv[k][a];
              crow[(size_t)a * rowlen + h] = cv[k][a];
            }
        }
        if (lane == 0 && wbk[k]) atomicOr(&gor[gi][jj[k]], wbk[k]);
        if (Q && MODE == kCfr) {
          const float wm = wave_max(lmx[k]);
          if (lane == 0) atomicMax(&gmx[gi][jj[k]], __float_as_int(wm));
        }
      }
    }
    __syncthreads();
    // the group's children: alive flags (every wave keeps its own copy of
    // the members' flags, the same values at the same LDS bytes, so a wave
    // reads flags it wrote itself), and the buffers two groups ahead zeroed
    for (uint32_t j = 0; j < gn; j++) {
      const MemRec& r = rec[ents[e + j] & 0xFFFF];
      if (!span.has(r.node)) continue;
      const uint32_t all = m_alive[r.src[p]] ? gor[gi][j] : 0u;
      if (lane < r.na) m_alive[r.clocal + lane] = (uint8_t)((all >> lane) & 1u);
      if (tid < r.na) alive[B.s(r.cslot[p]) + tid] = (all >> tid) & 1u;
    }
    const uint32_t gz = gi == 0 ? 2 : gi - 1;
    if (Q && MODE == kCfr) {
      // the new sums, encoded with each node's largest as its new scale
      // (storage.rs:306-320); every thread decoded with the old scales
      // before the barrier above
      for (uint32_t u = wave; u < units; u += nw) {
        const uint32_t j = u / tiles, h = (u - j * tiles) * 64 + lane;
        const MemRec& r = rec[ents[e + j] & 0xFFFF];
        // (a dead node was never staged: it keeps its sums and scale)
        if (!span.has(r.node) || !m_alive[r.src[p]] || h >= nq) continue;
        uint32_t yo = 0;
        for (uint32_t k = 0; k < j; k++) yo += rec[ents[e + k] & 0xFFFF].na;
        const float mx = __int_as_float(gmx[gi][j]);
        const float inv = mx > 0.0f ? kSumMax / mx : 0.0f;
        for (uint32_t a = 0; a < r.na; a++)
          sto.sq[r.offset + (size_t)a * nq + h] = mx > 0.0f ? enc_sum(ybuf[(yo + a) * nq + h], inv) : (uint16_t)0;
      }
      if (tid < gn) {
        const MemRec& r = rec[ents[e + tid] & 0xFFFF];
        if (span.has(r.node) && m_alive[r.src[p]]) sto.ssc[r.pslot] = __int_as_float(gmx[gi][tid]);
      }
      __syncthreads();  // the staging rows and gmx serve the next group
    }
    if (tid < DG_N) {
      gor[gz][tid] = 0;
      gmx[gz][tid] = 0;
    }
    gi = gi == 2 ? 0 : gi + 1;
    e += gn;
  }
  // every member's alive flag, for the up pass
  for (uint32_t m = tid; m < mcnt; m += blockDim.x) P.malive[moff + m] = m_alive[rec[m].src[p]];
}

// ------------------------------------------------------------------ terminals
// One wave per terminal. Sums over the opponent's reach are f64 chains run
// sequentially, one lane per chain, in the reference's order:
//   folds (eval.rs:40-64,172-206): total and per-card sums in hand order,
//     or, for a fold paired with a showdown, in strength order
//     (eval.rs:277-301);
//   showdowns (eval.rs:69-148,214-344): ascending and descending strength
//     sweeps, kept as prefix sums so every hand reads the running sums the
//     reference's sweep would hold at its strength.
// The values a chain reads are first gathered into LDS in chain order by the
// whole wave; each chain then runs over consecutive LDS words, in place.
// Sequential f64 chains over x[a..b), 16 values loaded ahead of the
// dependent adds (LDS latency overlaps the add chain).
template <bool STORE, bool FWD>
__device__ __forceinline__ double chain(double* x, uint32_t a, uint32_t b) {
  double acc = 0.0;
  const uint32_t n = b - a;
  uint32_t i = 0;
  auto at = [&](uint32_t k) -> double* { return FWD ? x + a + k : x + b - 1 - k; };
  if (n >= 16) {
    double cur[16];
#pragma unroll
    for (int k = 0; k < 16; k++) cur[k] = *at(k);
    for (; i + 32 <= n; i += 16) {
      double nxt[16];
#pragma unroll
      for (int k = 0; k < 16; k++) nxt[k] = *at(i + 16 + k);
#pragma unroll
      for (int k = 0; k < 16; k++) {
        acc += cur[k];
        if (STORE) *at(i + k) = acc;
      }
#pragma unroll
      for (int k = 0; k < 16; k++) cur[k] = nxt[k];
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
      acc += cur[k];
      if (STORE) *at(i + k) = acc;
    }
    i += 16;
  }
  for (; i < n; i++) {
    acc += *at(i);
    if (STORE) *at(i) = acc;
  }
  return acc;
}
__device__ __forceinline__ void chain_fwd(double* x, uint32_t a, uint32_t b) { chain<true, true>(x, a, b); }
__device__ __forceinline__ void chain_bwd(double* x, uint32_t a, uint32_t b) { chain<true, false>(x, a, b); }
__device__ __forceinline__ double chain_sum(double* x, uint32_t a, uint32_t b) { return chain<false, true>(x, a, b); }

// Gathers x[j] = (double)rv[idx[j]] for j < n, 8 index loads in flight per lane.
template <typename V>
__device__ __forceinline__ void gather(double* x, const V* rv, const uint16_t* __restrict__ idx, uint32_t n) {
  if (n == 0) return;
  for (uint32_t j0 = 0; j0 < n; j0 += TERM_T * 8) {
    uint32_t ix[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t j = min(j0 + k * TERM_T + threadIdx.x, n - 1);
      ix[k] = idx[j];
    }
    double v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = (double)rv[ix[k]];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t j = j0 + k * TERM_T + threadIdx.x;
      if (j < n) x[j] = v[k];
    }
  }
}

template <typename V>
__global__ void __launch_bounds__(TERM_T) terminal_kernel(DPlan P, int p, uint32_t t0, Span span, Bias B, const V* __restrict__ reach,
                                                          const uint32_t* __restrict__ alive, V* __restrict__ vals,
                                                          uint32_t rowlen) {
  extern __shared__ double sh[];
  __shared__ double tot[53];
  const TermRec tr = P.trec[t0 + blockIdx.x];
  if (!span.has(tr.node)) return;
  const uint32_t s = B.s(pick(tr.slot, p));
  if (!alive[s]) return;
  const int o = p ^ 1;
  const SpotTab& S = P.spots[tr.spot];
  const uint32_t nq = S.n[o], np = S.n[p], tid = threadIdx.x;
  double* Pa = sh;               // [rowlen]
  double* Ca = sh + rowlen;      // [2 * rowlen]
  V* rv = (V*)(sh + 3 * rowlen); // [rowlen]: the opponent's reach row, hand order
  const V* R = reach + (size_t)s * rowlen;
  V* out = vals + (size_t)B.v(tr.vslot) * rowlen;
  for (uint32_t j0 = 0; j0 < nq; j0 += TERM_T * 8) {
    V x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t j = j0 + k * TERM_T + tid;
      x[k] = j < nq ? R[j] : V(0);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint32_t j = j0 + k * TERM_T + tid;
      if (j < nq) rv[j] = x[k];
    }
  }
  __syncthreads();
  const bool batched = (tr.tflags & kTermBatched) != 0 && !span.kid(tr.node);
  const double hp = S.half_pot, norm = S.norm, rake = (double)tr.rake;
  const int32_t inv_p = pick(tr.invest, p), inv_o = pick(tr.invest, o);
  const uint8_t* cp = S.cards[p];
  const uint8_t* co = S.cards[o];
  const uint32_t* same_p = S.same[p];

  if (tr.kind != kShowdown) {
    const int f = tr.kind - kFold0;
    // cfr.rs:224-235
    const double pay = (f == p ? -(hp + (double)inv_p) : hp + (double)pick(tr.invest, f) - rake) * norm;
    const bool paired = batched && (tr.tflags & (p == 0 ? kTermPaired0 : kTermPaired1)) != 0;
    if (paired) {  // strength order of the sibling showdown's board
      const uint32_t b = tr.board;
      const uint32_t cnt = S.ent_cnt[o][b];
      const uint16_t* clo = S.cl_off[o] + (size_t)b * 53;
      gather(Pa, rv, S.ent_idx[o] + S.ent_off[o][b], cnt);
      gather(Ca, rv, S.cl_idx[o] + S.cl_base[o][b], 2 * cnt);
      __syncthreads();
      if (tid == 52) tot[52] = chain_sum(Pa, 0, cnt);
      else if (tid < 52) tot[tid] = chain_sum(Ca, clo[tid], clo[tid + 1]);
    } else {  // hand order
      const uint32_t* bo = S.bl_off[o];
      gather(Ca, rv, S.bl_idx[o], bo[52]);
      __syncthreads();
      for (uint32_t j = tid; j < nq; j += TERM_T) Pa[j] = (double)rv[j];
      __syncthreads();
      if (tid == 52) {
        tot[52] = chain_sum(Pa, 0, nq);
      } else if (tid < 52) {
        tot[tid] = chain_sum(Ca, bo[tid], bo[tid + 1]);
      }
    }
    __syncthreads();
    const double total = tot[52];
    const uint16_t* cp2 = (const uint16_t*)cp;
    for (uint32_t i0 = 0; i0 < np; i0 += TERM_T * 8) {
      uint32_t cc[8], sm[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t i = min(i0 + k * TERM_T + tid, np - 1);
        cc[k] = cp2[i];
        sm[k] = same_p[i];
      }
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t i = i0 + k * TERM_T + tid;
        const int a = cc[k] & 0xFF, b = cc[k] >> 8;
        const double sw = sm[k] == NONE ? 0.0 : (double)rv[sm[k]];
        V v;
        if (batched) v = (V)(pay * (total - tot[a] - tot[b] + sw));
        else v = total == 0.0 ? V(0) : (V)(pay * (total - (tot[a] + tot[b]) + sw));
        if (i < np) out[i] = v;
      }
    }
    return;
  }

  // showdown (cfr.rs:238-246 payoffs)
  const double win = (hp + (double)inv_o - rake) * norm;
  const double lose = -(hp + (double)inv_p) * norm;
  const double tie = 0.5 * ((double)(inv_o - inv_p) - rake) * norm;
  const uint32_t b = tr.board;
  const uint32_t mo = S.ent_cnt[o][b];
  const uint16_t* eidx = S.ent_idx[o] + S.ent_off[o][b];
  const uint16_t* clo = S.cl_off[o] + (size_t)b * 53;
  const uint16_t* ci = S.cl_idx[o] + S.cl_base[o][b];
  const uint4* hcnt = S.hand_cnt[p] + (size_t)b * np;
  const uint16_t* cp2 = (const uint16_t*)cp;

  // ascending sweep: weaker opponent hands
  gather(Pa, rv, eidx, mo);
  gather(Ca, rv, ci, 2 * mo);
  __syncthreads();
  if (tid == 52) chain_fwd(Pa, 0, mo);
  else if (tid < 52) chain_fwd(Ca, clo[tid], clo[tid + 1]);
  __syncthreads();
  // per-hand counts and cards for this board, loaded once (clamped indices:
  // no load waits behind a branch)
  uint4 hc[MAX_K];
  uint32_t hcards[MAX_K];
#pragma unroll
  for (int kk = 0; kk < MAX_K; kk++) {
    const uint32_t i = min(tid + kk * TERM_T, np - 1);
    hc[kk] = hcnt[i];
    hcards[kk] = cp2[i];
  }
  double W[MAX_K];
#pragma unroll
  for (int kk = 0; kk < MAX_K; kk++) {
    const uint32_t lt = hc[kk].x & 0xFFFF;
    const uint32_t l1 = hc[kk].y & 0xFFFF, l2 = hc[kk].z & 0xFFFF;
    const int c1 = hcards[kk] & 0xFF, c2 = hcards[kk] >> 8;
    const double Pv = (lt != 0xFFFF && lt) ? Pa[lt - 1] : 0.0;
    const double P1 = l1 ? Ca[clo[c1] + l1 - 1] : 0.0;
    const double P2 = l2 ? Ca[clo[c2] + l2 - 1] : 0.0;
    W[kk] = batched ? win * (Pv - P1 - P2) : win * (Pv - (P1 + P2));
  }
  __syncthreads();
  // descending sweep: stronger opponent hands
  gather(Pa, rv, eidx, mo);
  gather(Ca, rv, ci, 2 * mo);
  __syncthreads();
  if (tid == 52) chain_bwd(Pa, 0, mo);
  else if (tid < 52) chain_bwd(Ca, clo[tid], clo[tid + 1]);
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < MAX_K; kk++) {
    const uint32_t i = tid + kk * TERM_T;
    const uint32_t lt = hc[kk].x & 0xFFFF, le = hc[kk].x >> 16;
    const uint32_t m1 = hc[kk].y >> 16, m2 = hc[kk].z >> 16;
    const int c1 = hcards[kk] & 0xFF, c2 = hcards[kk] >> 8;
    V v = V(0);
    if (lt != 0xFFFF) {
      const uint32_t len1 = clo[c1 + 1] - clo[c1], len2 = clo[c2 + 1] - clo[c2];
      const double Sv = le < mo ? Pa[le] : 0.0;
      const double S1 = m1 < len1 ? Ca[clo[c1] + m1] : 0.0;
      const double S2 = m2 < len2 ? Ca[clo[c2] + m2] : 0.0;
      if (batched) {
        const double r = W[kk] + lose * (Sv - S1 - S2);
        v = (V)r;
      } else {
        v = (V)W[kk];
        v = v + (V)(lose * (Sv - (S1 + S2)));
      }
      if (tie != 0.0 && i < np) {  // eval.rs:347-379
        double es = 0.0, e1 = 0.0, e2 = 0.0;
        for (uint32_t j = lt; j < le; j++) {
          const uint32_t idx = eidx[j];
          const double w = (double)rv[idx];
          es += w;
          if (holds(co, idx, c1)) e1 += w;
          if (holds(co, idx, c2)) e2 += w;
        }
        const uint32_t sm = same_p[i];
        const double sw = sm == NONE ? 0.0 : (double)rv[sm];
        const double eq = es - (e1 + e2) + sw;
        v = v + (V)(tie * eq);
      }
    }
    if (i < np) out[i] = v;
  }
}

// ------------------------------------------------------------------ fast terminals
// The same terminal values with every f64 sum computed as a parallel scan
// instead of one lane's sequential chain: the sums differ from the
// reference's only in summation order (a few f64 ulps), so an f32 value
// occasionally differs by one rounding step, as the reference's own `Fast`
// precision does (crates/solver/src/gpu/mod.rs:51-63).
//
// Only prefix sums are stored, as the reference's GPU showdown does
// (crates/solver/cuda/kernels/terminal.cu:307-384): a suffix is the total
// less a prefix. A suffix whose terms are all zero is set to an exact zero
// (the last nonzero strength position is kept per terminal), so a hand
// nothing beats sums to zero as in the sequential order.
//
// One workgroup per run of terminals that share a river board (TermBatch).
// Per run, every lane computes its LDS addresses once: the strength-order
// and card-list entries it gathers, and per hand the prefix slots it reads.
// Per terminal:
//   the reach row to LDS (skipped when the previous terminal read the same
//   row); barrier;
//   waves 0-1: the opponent's strength-order exclusive prefix (11 entries
//   per lane, wave scans of the lane totals; wave 1's prefixes exclude wave
//   0's total, added at the read); waves 2-5: per-card exclusive prefixes,
//   4 lanes per card, 13 positions each, at 52 slots per card, and the
//   card totals; barrier;
//   one value per hand; barrier.
// A fold reads only totals (its scans stop there), and a fold that follows
// its sibling showdown on the same reach row (river.rs:163-213 pairs them)
// reuses that showdown's sums.
constexpr int FT = 512;                   // fast terminal workgroup
constexpr int FK = (1326 + FT - 1) / FT;  // hands per lane (strided)
constexpr int SLN = 128;                  // strength-order lanes (waves 0 and 1)
constexpr int SE = 11;                    // strength-order entries per lane
constexpr uint32_t SH = 64 * SE;          // wave 0's entries
constexpr int CG = 4;                     // lanes per card list
constexpr int CE = 13;                    // card-list positions per lane
constexpr int CL0 = SLN;                  // first card lane
constexpr int CR = CG * CE;               // prefix slots per card
constexpr uint32_t ZR = 1344;             // the reach row's zero slot (rows hold at most 1326 hands)
static_assert(SLN * SE >= 1326 + 1, "strength prefixes cover every position of a 1326-hand list");
static_assert(CR >= 51 + 1, "card prefixes cover every position of a 51-hand list");
static_assert(CL0 + 52 * CG <= FT && CL0 % 64 == 0, "card lanes fit the workgroup, groups inside waves");
static_assert(ZR >= ((1326 + 31) & ~31), "the zero slot follows the longest row");

template <typename V>
struct TermLds {
  V rv[ZR + 1];           // the opponent's reach row, hand order; rv[ZR] = 0
  double pa[SLN * SE];    // strength-order exclusive prefixes (wave 1's without wave 0's total)
  double ca[52 * CR];     // per card: exclusive prefixes of its list
  double ct[52];          // per card: its list's total
  double wt[2];           // the strength-order totals of waves 0 and 1
  uint32_t lz[2];         // per wave: one past the last nonzero strength position (0: none)
};

// The fields of a terminal record one traversal of player p reads, as
// wave-uniform scalars. Records and alive flags are read through the
// constant address space, so they are scalar loads counted by lgkmcnt: a
// vector load here would make each terminal wait (vmcnt) for the value-row
// stores of the one before it. (Both arrays are read-only in this kernel.)
typedef const __attribute__((address_space(4))) uint32_t* cu32p;
struct TermFields {
  uint32_t node, slot, vslot, kf;  // kf: kind | tflags << 8
  int32_t inv0, inv1;
  float rake;
};
template <int p>
__device__ __forceinline__ TermFields term_fields(const TermRec* __restrict__ r) {
  cu32p w = (cu32p)r;  // node, slot[2], vslot, board, invest[2], rake, kind | tflags << 8
  TermFields f;
  f.node = w[0];
  f.slot = w[1 + p];
  f.vslot = w[3];
  f.inv0 = (int32_t)w[5];
  f.inv1 = (int32_t)w[6];
  f.rake = __uint_as_float(w[7]);
  f.kf = w[8] & 0xFFFFu;
  return f;
}
static_assert(offsetof(TermRec, invest) == 20 && offsetof(TermRec, rake) == 28 && offsetof(TermRec, kind) == 32 &&
                  offsetof(TermRec, tflags) == 33,
              "TermRec word layout");

// f64 lane moves by DPP (two 32-bit moves): lane i reads the lane CTRL
// names; lanes without a source (or in rows ROW_MASK leaves out) read 0.
// row_shr:n 0x110 + n, row_shl:n 0x100 + n (within rows of 16 lanes);
// wave_shr:1 0x138, wave_shl:1 0x130; row_bcast:15 0x142, row_bcast:31 0x143.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW_MASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW_MASK, 0xF, false);
  return __hiloint2double(hi, lo);
}
// The wave's largest u32 (0 is the identity), in every lane: the steps of wave_max.
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false));
  return (uint32_t)__builtin_amdgcn_readlane((int)x, 63);
}

// six waves per SIMD in both precisions (the f64 form fits 78 VGPRs there without
// spilling: flop FP64 check 159.4 -> 148.1 ms against five)
template <typename V, int p>  // p: the traverser (a constant, so per-player tables are static fields)
__global__ void __launch_bounds__(FT, 6) terminal_fast_kernel(DPlan P, uint32_t b0, Span span, Bias B,
                                                              const V* __restrict__ reach, const uint32_t* __restrict__ alive,
                                                              V* __restrict__ vals, uint32_t rowlen) {
  __shared__ TermLds<V> L;
  const TermBatch tb = P.tbatch[b0 + blockIdx.x];
  constexpr int o = p ^ 1;
  const SpotTab& S = P.spots[tb.spot];
  const uint32_t nq = S.n[o], np = S.n[p], tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t b = tb.board;
  const bool sorted = b != NONE;  // strength order of the board, else hand order (folds only)

  // Per hand (strided): h0 = lt | le << 16 against the opponent's strength
  // order (lt 0xFFFF: the board blocks the hand); h1, h2 = the card-prefix
  // slots of its first and second card at lt(card) | le(card) << 16; h3 =
  // c1 | c2 << 8 | the reach slot of the identical opponent combo << 16 (ZR: none).
  uint32_t h0[FK], h1[FK], h2[FK], h3[FK];
  {
    const uint16_t* cp2 = (const uint16_t*)S.cards[p];
    const uint4* hcnt = S.hand_cnt[p] + (size_t)(sorted ? b : 0) * np;
#pragma unroll
    for (int k = 0; k < FK; k++) {
      const uint32_t i = tid + k * FT;
      uint32_t c1 = 0, c2 = 0, sm = ZR;
      uint4 hc = make_uint4(0xFFFFu, 0, 0, 0);
      if (i < np) {
        const uint32_t cc = cp2[i];
        c1 = cc & 0xFF;
        c2 = cc >> 8;
        const uint32_t s = S.same[p][i];
        sm = s == NONE ? ZR : s;
        if (sorted) hc = hcnt[i];
      }
      h0[k] = hc.x;
      h1[k] = (c1 * CR + (hc.y & 0xFFFF)) | (c1 * CR + (hc.y >> 16)) << 16;
      h2[k] = (c2 * CR + (hc.z & 0xFFFF)) | (c2 * CR + (hc.z >> 16)) << 16;
      h3[k] = c1 | c2 << 8 | sm << 16;
    }
  }
  // The entries this lane gathers (reach slots, ZR past the list): waves 0-1
  // a contiguous slice of the strength order; card lanes a slice of their
  // card's list.
  uint32_t src[(CE + 1) / 2];  // two 16-bit reach slots per register
  const bool slane = tid < SLN, clane = tid >= CL0 && tid < CL0 + 52 * CG;
  const uint32_t card = clane ? (tid - CL0) / CG : 0, cq = (tid - CL0) & (CG - 1);
  {
    const uint32_t nA = sorted ? S.ent_cnt[o][b] : nq;
    uint32_t s1[CE];
#pragma unroll
    for (int k = 0; k < CE; k++) s1[k] = ZR;
    if (slane) {
      const uint16_t* idxA = S.ent_idx[o] + (sorted ? S.ent_off[o][b] : 0u);
#pragma unroll
      for (int k = 0; k < SE; k++) {
        const uint32_t j = tid * SE + k;
        if (j < nA) s1[k] = sorted ? (uint32_t)idxA[j] : j;
      }
    } else if (clane) {
      const uint32_t a = sorted ? (uint32_t)S.cl_off[o][(size_t)b * 53 + card] : S.bl_off[o][card];
      const uint32_t e = sorted ? (uint32_t)S.cl_off[o][(size_t)b * 53 + card + 1] : S.bl_off[o][card + 1];
      const uint16_t* idx = sorted ? S.cl_idx[o] + S.cl_base[o][b] : S.bl_idx[o];
#pragma unroll
      for (int k = 0; k < CE; k++) {
        const uint32_t x = a + cq * CE + k;
        if (x < e) s1[k] = idx[x];
      }
    }
#pragma unroll
    for (int k = 0; k < CE; k += 2) src[k / 2] = s1[k] | (k + 1 < CE ? s1[k + 1] : ZR) << 16;
  }
  auto slot_of = [&](int k) -> uint32_t { return (k & 1) ? src[k / 2] >> 16 : src[k / 2] & 0xFFFF; };
  if (tid == 0) L.rv[ZR] = V(0);
  const double hp = S.half_pot, norm = S.norm;

  // the first terminal's record, flag and reach row
  uint32_t t = tb.t0;
  const uint32_t tend = tb.t0 + tb.cnt;
  TermFields tr = term_fields<p>(P.trec + t);
  uint32_t live = span.has(tr.node) ? ((cu32p)alive)[B.s(tr.slot)] : 0u;
  V r[FK];
  {
    const V* R = reach + (size_t)B.s(tr.slot) * rowlen;
#pragma unroll
    for (int k = 0; k < FK; k++) {
      const uint32_t j = tid + k * FT;
      r[k] = j < nq ? R[j] : V(0);
    }
  }
  uint32_t have_slot = NONE;  // the slot whose row is in L.rv
  int have = 0;               // what L holds for it: 0 nothing, 1 totals, 2 prefixes and totals
  for (;;) {
    // The per-lane tables are re-declared live each terminal (empty asm), so
    // what is derived from them (addresses, range tests) is recomputed here
    // instead of being hoisted out of the loop into registers that spill.
#pragma unroll
    for (int k = 0; k < FK; k++) asm volatile("" : "+v"(h0[k]), "+v"(h1[k]), "+v"(h2[k]), "+v"(h3[k]));
#pragma unroll
    for (int k = 0; k < (CE + 1) / 2; k++) asm volatile("" : "+v"(src[k]));
    // fetch the next terminal while this one is summed (its slot is in
    // range whether or not it is alive); a terminal on the same row reuses it
    const bool more = t + 1 < tend;
    TermFields trn = tr;
    uint32_t liven = 0;
    V rn[FK];
#pragma unroll
    for (int k = 0; k < FK; k++) rn[k] = r[k];
    if (more) {
      trn = term_fields<p>(P.trec + t + 1);
      const uint32_t sn = B.s(trn.slot);
      liven = span.has(trn.node) ? ((cu32p)alive)[sn] : 0u;
      if (trn.slot != tr.slot) {
        const V* R = reach + (size_t)sn * rowlen;
#pragma unroll
        for (int k = 0; k < FK; k++) {
          const uint32_t j = tid + k * FT;
          rn[k] = j < nq ? R[j] : V(0);
        }
      }
    }
    if (live) {  // a dead slot's value row is never read
      const uint32_t kind = tr.kf & 0xFF;
      const bool showdown = kind == kShowdown;
      const int need = showdown ? 2 : 1;
      if (tr.slot != have_slot) {
#pragma unroll
        for (int k = 0; k < FK; k++) {
          const uint32_t j = tid + k * FT;
          if (j < nq) L.rv[j] = r[k];
        }
        have_slot = tr.slot;
        have = 0;
        __syncthreads();
      }
      if (have < need) {
        if (slane) {
          V v[SE];
          double tot = 0.0;
          uint32_t last = 0;
#pragma unroll
          for (int k = 0; k < SE; k++) v[k] = L.rv[slot_of(k)];
#pragma unroll
          for (int k = 0; k < SE; k++) {
            tot += (double)v[k];
            if (v[k] != V(0)) last = tid * SE + k + 1;
          }
          // inclusive scan of the lane totals across the wave: DPP row
          // shifts within rows of 16, then rows 0 and 2's last lanes into
          // rows 1 and 3, then row 1's last lane into rows 2 and 3
          double inc = tot;
          inc += dpp_f64<0x111>(inc);
          inc += dpp_f64<0x112>(inc);
          inc += dpp_f64<0x114>(inc);
          inc += dpp_f64<0x118>(inc);
          inc += dpp_f64<0x142, 0xA>(inc);
          inc += dpp_f64<0x143, 0xC>(inc);
          if (need == 2) {
            double run = dpp_f64<0x138>(inc);  // wave_shr:1
            if (lane == 0) run = 0.0;
#pragma unroll
            for (int k = 0; k < SE; k++) {
              L.pa[tid * SE + k] = run;
              run += (double)v[k];
            }
            last = wave_max_u32(last);
          }
          if (lane == 63) {
            L.wt[wave] = inc;
            L.lz[wave] = last;
          }
        } else if (clane) {
          V v[CE];
          double tot = 0.0;
#pragma unroll
          for (int k = 0; k < CE; k++) v[k] = L.rv[slot_of(k)];
#pragma unroll
          for (int k = 0; k < CE; k++) tot += (double)v[k];
          // inclusive scan over the card's 4 lanes (a value from the
          // neighbouring group is masked out)
          double inc = tot;
          const double y1 = dpp_f64<0x111>(inc);
          if (cq >= 1) inc += y1;
          const double y2 = dpp_f64<0x112>(inc);
          if (cq >= 2) inc += y2;
          if (need == 2) {
            double run = dpp_f64<0x111>(inc);
            if (cq == 0) run = 0.0;
#pragma unroll
            for (int k = 0; k < CE; k++) {
              L.ca[card * CR + cq * CE + k] = run;
              run += (double)v[k];
            }
          }
          if (cq == CG - 1) L.ct[card] = inc;
        }
        have = need;
        __syncthreads();
      }
      const double rake = (double)tr.rake;
      const int32_t inv_p = p ? tr.inv1 : tr.inv0, inv_o = p ? tr.inv0 : tr.inv1;
      const bool batched = ((tr.kf >> 8) & kTermBatched) != 0 && !span.kid(tr.node);
      V* out = vals + (size_t)B.v(tr.vslot) * rowlen;
      const double W0 = L.wt[0], total = W0 + L.wt[1];
      if (!showdown) {  // cfr.rs:224-235, eval.rs:39-64,170-206
        const int f = (int)kind - kFold0;
        const double pay = (f == p ? -(hp + (double)inv_p) : hp + (double)(f ? tr.inv1 : tr.inv0) - rake) * norm;
#pragma unroll
        for (int k = 0; k < FK; k++) {
          const uint32_t i = tid + k * FT;
          if (i >= np) continue;
          const double ta = L.ct[h3[k] & 0xFF], tc = L.ct[(h3[k] >> 8) & 0xFF];
          const double sw = (double)L.rv[h3[k] >> 16];
          V x;
          if (batched) x = (V)(pay * (total - ta - tc + sw));
          else x = total == 0.0 ? V(0) : (V)(pay * (total - (ta + tc) + sw));
          out[i] = x;
        }
      } else {  // showdown: cfr.rs:238-246 payoffs, eval.rs:66-148,208-379 sweeps
        const uint32_t lz = max(L.lz[0], L.lz[1]);  // one past the last nonzero strength position
        const double win = (hp + (double)inv_o - rake) * norm;
        const double lose = -(hp + (double)inv_p) * norm;
        const double tie = 0.5 * ((double)(inv_o - inv_p) - rake) * norm;
#pragma unroll
        for (int k = 0; k < FK; k++) {
          const uint32_t i = tid + k * FT;
          if (i >= np) continue;
          const uint32_t lt = h0[k] & 0xFFFF, le = h0[k] >> 16;
          V x = V(0);
          if (lt != 0xFFFF) {
            const double Pv = L.pa[lt] + (lt >= SH ? W0 : 0.0);
            const double Pe = L.pa[le] + (le >= SH ? W0 : 0.0);
            const double P1 = L.ca[h1[k] & 0xFFFF], Q1 = L.ca[h1[k] >> 16];
            const double P2 = L.ca[h2[k] & 0xFFFF], Q2 = L.ca[h2[k] >> 16];
            double Sv = 0.0, S1 = 0.0, S2 = 0.0;
            if (le < lz) {  // something stronger has reach (else every stronger sum is an exact zero)
              Sv = total - Pe;
              S1 = L.ct[h3[k] & 0xFF] - Q1;
              S2 = L.ct[(h3[k] >> 8) & 0xFF] - Q2;
            }
            if (batched) {
              x = (V)(win * (Pv - P1 - P2) + lose * (Sv - S1 - S2));
            } else {
              x = (V)(win * (Pv - (P1 + P2)));
              x = x + (V)(lose * (Sv - (S1 + S2)));
            }
            if (tie != 0.0) {  // eval.rs:347-379: the tied block, less the hands holding a card, plus the same combo
              const double sw = (double)L.rv[h3[k] >> 16];
              x = x + (V)(tie * ((Pe - Pv) - ((Q1 - P1) + (Q2 - P2)) + sw));
            }
          }
          out[i] = x;
        }
      }
      // the next terminal rewrites L unless it reads this row's sums as they are
      const bool keep = more && liven && trn.slot == tr.slot && ((trn.kf & 0xFF) != kShowdown || have == 2);
      if (!keep) __syncthreads();
    }
    if (!more) break;
    t++;
    tr = trn;
    live = liven;
#pragma unroll
    for (int k = 0; k < FK; k++) r[k] = rn[k];
  }
}

// ------------------------------------------------------------------ chance
// Chance nodes of one segment depth, after the deeper segments' up pass:
// the f64 sum over dealt cards of the child rows, zero for own hands holding
// the card (cfr.rs:457-478); dead nodes get a zero row, and a dead child adds
// an exact zero (its row is not written when it is one terminal or chance
// node). One lane per hand, 16 child rows in flight per lane.
template <typename V>
__global__ void __launch_bounds__(UP_T) chance_kernel(DPlan P, int p, uint32_t c0, Span span, Bias B, V* __restrict__ vals,
                                                      uint32_t rowlen, const uint32_t* __restrict__ alive) {
  const uint32_t ci = c0 + blockIdx.x;
  const MemRec& r = P.crec[ci];
  const SpotTab& S = P.spots[r.spot];
  const uint32_t np = S.n[p];
  const uint32_t h = blockIdx.y * UP_T + threadIdx.x;
  if (!span.has(r.node) || h >= np) return;
  V* out = vals + (size_t)B.v(r.vslot) * rowlen + h;
  if (!alive[B.s(r.slot[p])]) {
    *out = V(0);
    return;
  }
  const uint32_t na = r.na;
  const V* row = vals + (size_t)B.v(P.chance_child_vslot[ci]) * rowlen + h;
  const uint32_t* calive = alive + B.s(r.cslot[p]);  // the children's alive flags (consecutive slots)
  const int8_t* cards = P.card + r.offset;
  const int c1 = S.cards[p][2 * h], c2 = S.cards[p][2 * h + 1];
  double acc = 0.0;
  uint32_t a = 0;
  for (; a + 16 <= na; a += 16) {
    V x[16];
#pragma unroll
    for (int k = 0; k < 16; k++) x[k] = calive[a + k] ? row[(size_t)(a + k) * rowlen] : V(0);
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int c = cards[a + k];
      acc += (c == c1 || c == c2) ? 0.0 : (double)x[k];
    }
  }
  for (; a < na; a++) {
    const int c = cards[a];
    const V x = calive[a] ? row[(size_t)a * rowlen] : V(0);
    acc += (c == c1 || c == c2) ? 0.0 : (double)x;
  }
  // cards merged by suit isomorphism, after the canonical ones: the
  // canonical child's row at the permuted hand (cfr.rs:457-478)
  const V* base = row - h;
  for (uint32_t i = P.ciso_off[ci]; i < P.ciso_off[ci + 1]; i++) {
    const IsoRef e = P.ciso[i];
    const uint32_t hp = S.perm[p][(size_t)e.pair * np + h];
    const V x = calive[e.child] ? base[(size_t)e.child * rowlen + hp] : V(0);
    acc += ((int)e.card == c1 || (int)e.card == c2) ? 0.0 : (double)x;
  }
  *out = (V)acc;
}

// ------------------------------------------------------------------ up
// f32 walks, and the f64 best-response walk, keep a segment's member values
// in VGPRs (up_reg_kernel<V, MODE, CAP>: up to 32 members in one vector
// register range, up to 48 at four waves per SIMD, up to 64 and 80 at two);
// larger segments, and the other f64 walks, keep them in LDS (up_kernel).
constexpr uint32_t REG_T = 32, REG_S = 48, REG_M = 64, REG_L = 80;
constexpr uint64_t kNoRow = ~0ull;

// CAP 32-bit values (T: f32 or u32) per lane in up to three vector registers (32 + 32 + CAP - 64 lanes of
// the member index), read and written at a wave-uniform index (indexed
// register moves: no LDS, no scratch). The vectors are separate local
// variables of the kernel (each small enough for the compiler to keep in
// registers under a variable index); indices are masked into range and the
// selects take values, not addresses.
template <int CAP, typename T = float>
struct MemberVec {
  static_assert(CAP == 32 || CAP == 48 || CAP == 64 || CAP == 80, "32, 48, 64 or 80 members");
  static constexpr int W1 = CAP == 32 ? 2 : CAP >= 64 ? 32 : CAP - 32;  // (CAP 32: an unused pair)
  static constexpr int W2 = CAP > 64 ? CAP - 64 : 2;                   // (CAP <= 64: an unused pair)
  typedef T lo __attribute__((ext_vector_type(32)));
  typedef T mi __attribute__((ext_vector_type(W1)));
  typedef T hi __attribute__((ext_vector_type(W2)));
  __device__ static __forceinline__ T get(const lo& r0, const mi& r1, const hi& r2, uint32_t i) {
    const T x0 = r0[i & 31];
    if (CAP == 32) return x0;
    const T x1 = r1[i & (W1 - 1)];
    if (CAP <= 64) return i < 32 ? x0 : x1;
    const T x2 = r2[i & (W2 - 1)];
    return i < 32 ? x0 : i < 64 ? x1 : x2;
  }
  // Every vector is written (its element kept where i is not in it), with
  // no branch on i: a branch would make the compiler merge the vectors of
  // both paths by copying a whole vector per write.
  __device__ static __forceinline__ void set(lo& r0, mi& r1, hi& r2, uint32_t i, T v) {
    if (CAP == 32) {
      r0[i & 31] = v;
      return;
    }
    const T o0 = r0[i & 31];
    r0[i & 31] = i < 32 ? v : o0;
    const T o1 = r1[i & (W1 - 1)];
    r1[i & (W1 - 1)] = (i >= 32 && (CAP <= 64 || i < 64)) ? v : o1;
    if (CAP > 64) {
      const T o2 = r2[i & (W2 - 1)];
      r2[i & (W2 - 1)] = i >= 64 ? v : o2;
    }
  }
};

// Member values of type V: f32 in one MemberVec (f0, f1, f2); f64 as its
// low and high words in two u32 ones (w0.. and x0..): a variable index
// into a vector of 64-bit elements is lowered through scratch memory, into
// 32-bit ones by indexed register moves. The pairs V does not use are dead
// and compiled away. (Separate locals, not a struct: a struct of vectors
// under a variable index is kept in scratch memory.)
template <typename V, int CAP>
__device__ __forceinline__ V mval_get(const typename MemberVec<CAP>::lo& f0, const typename MemberVec<CAP>::mi& f1,
                                      const typename MemberVec<CAP>::hi& f2,
                                      const typename MemberVec<CAP, uint32_t>::lo& w0,
                                      const typename MemberVec<CAP, uint32_t>::mi& w1,
                                      const typename MemberVec<CAP, uint32_t>::hi& w2,
                                      const typename MemberVec<CAP, uint32_t>::lo& x0,
                                      const typename MemberVec<CAP, uint32_t>::mi& x1,
                                      const typename MemberVec<CAP, uint32_t>::hi& x2, uint32_t i) {
  typedef MemberVec<CAP, uint32_t> W;
  if constexpr (sizeof(V) == 4) return MemberVec<CAP>::get(f0, f1, f2, i);
  else return __hiloint2double((int)W::get(x0, x1, x2, i), (int)W::get(w0, w1, w2, i));
}
template <typename V, int CAP>
__device__ __forceinline__ void mval_set(typename MemberVec<CAP>::lo& f0, typename MemberVec<CAP>::mi& f1,
                                         typename MemberVec<CAP>::hi& f2, typename MemberVec<CAP, uint32_t>::lo& w0,
                                         typename MemberVec<CAP, uint32_t>::mi& w1, typename MemberVec<CAP, uint32_t>::hi& w2,
                                         typename MemberVec<CAP, uint32_t>::lo& x0, typename MemberVec<CAP, uint32_t>::mi& x1,
                                         typename MemberVec<CAP, uint32_t>::hi& x2, uint32_t i, V v) {
  typedef MemberVec<CAP, uint32_t> W;
  if constexpr (sizeof(V) == 4) {
    MemberVec<CAP>::set(f0, f1, f2, i, v);
  } else {
    W::set(w0, w1, w2, i, (uint32_t)__double2loint(v));
    W::set(x0, x1, x2, i, (uint32_t)__double2hiint(v));
  }
}

__device__ __forceinline__ uint32_t lane_u32(uint32_t x, uint32_t m) { return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)m); }
__device__ __forceinline__ float lane_f32(float x, uint32_t m) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), (int)m)); }
__device__ __forceinline__ uint64_t lane_u64(uint64_t x, uint32_t m) {
  return (uint64_t)lane_u32((uint32_t)x, m) | (uint64_t)lane_u32((uint32_t)(x >> 32), m) << 32;
}

// Record flags of a member held by its lane
constexpr uint32_t kFLive = 8, kFSpan = 16, kFRoot = 32, kFFrozen = 64;

// One wave per (segment, 64-hand tile), one hand per lane, the segment's
// members in registers: lane m holds member m's record, read by every lane
// with readlane; every value row the walk needs (terminal and chance rows,
// and in CFR / average modes the regret or strategy-sum row of each live own
// node's action, kept at the index of the child it leads to) is loaded
// before the walk, all loads in flight together; the walk then runs on
// registers, in reverse breadth-first order, with the arithmetic of
// up_kernel below. CAP = 32 (or 48) runs over a range of segments and leaves
// those above CAP members; CAP = 48, 64 and 80 also run over the plan's lists
// of segments of 33 to 48, 49 to 64 and 65 to 80 members (list[seg0 +
// blockIdx.x]). With CAP 80 each lane holds two records, members lane and
// lane + 64.
//
// QP, the storage pass: 0 f32 storage; with 16-bit storage 3 reads (average
// mode), and a CFR walk runs twice, 1 for the values and every own node's
// largest |new regret| (atomically maxed over the hand tiles into rnx), 2 for
// the same regrets encoded with that maximum as the node's new scale
// (storage.rs:253-289: the scale is the whole block's maximum).
//
// V: the value type, f32, or f64 in best-response mode (no storage rows).
template <typename V, int MODE, int CAP, int QP>
__device__ __forceinline__ void up_reg_walk(const DPlan& P, int p, uint32_t seg0, const Span& span, const Bias& B,
                                            V* __restrict__ vals, uint32_t rowlen, const uint32_t* __restrict__ alive,
                                            const Store& sto, const IterParams* __restrict__ prm, V* __restrict__ qout,
                                            const uint32_t* __restrict__ list) {
  static_assert(std::is_same<V, float>::value || MODE == kBest, "f64 walks in registers: best response only");
  constexpr bool ROWS = MODE != kBest;
  constexpr bool Q = QP != 0;
  __shared__ uint64_t src_lds[CAP];
  __shared__ float stp_lds[Q ? CAP : 1];  // 16-bit storage: the decoding step of each storage row, by child
  const uint32_t seg = list != nullptr ? list[seg0 + blockIdx.x] : seg0 + blockIdx.x;
  const uint4 head = P.seg_head[seg];
  const uint32_t moff = head.x, mcnt = head.y, root = head.z;
  if (mcnt > (uint32_t)CAP) return;
  if (!span.has(root) && seg != span.seg) return;
  const uint32_t np = p ? head.w >> 16 : head.w & 0xFFFF, lane = threadIdx.x;
  const uint32_t h = blockIdx.y * UP_T + lane;
  const bool act = h < np;
  const uint32_t hh = act ? h : 0;  // idle lanes load row element 0 and store nothing
  // member lane's record, and with CAP above 64 member lane + 64's as well
  constexpr int H = CAP > 64 ? 2 : 1;
  uint32_t f[H], vslot[H], psr[H];
  uint64_t off[H], src[H];
  float stp[H];
#pragma unroll
  for (int q = 0; q < H; q++) {
    const uint32_t mm = lane + 64 * q;
    f[q] = vslot[q] = psr[q] = 0;
    off[q] = 0;
    src[q] = kNoRow;
    stp[q] = 0.0f;
    if (mm < mcnt) {
      const MemRec& r = P.mrec[moff + mm];
      const bool lm = P.malive[moff + mm] != 0;
      f[q] = (uint32_t)r.kind | (lm ? kFLive : 0u) | (span.has(r.node) ? kFSpan : 0u) | (r.node == span.root ? kFRoot : 0u) |
             ((r.tflags & kNodeFrozen) ? kFFrozen : 0u) | (uint32_t)r.na << 8 | (uint32_t)r.clocal << 16;
      vslot[q] = r.vslot;
      off[q] = r.offset;
      psr[q] = r.pslot;
      if (ROWS) src_lds[mm] = kNoRow;
    }
  }
  if (ROWS) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < H; q++) {
      const uint32_t fq = f[q];
      if (lane + 64 * q < mcnt && (fq & (kFLive | kFSpan)) == (kFLive | kFSpan) && (int)(fq & 7) == p) {
        float step = 0.0f;
        if (Q) step = MODE == kCfr ? sto.rsc[psr[q]] / kRegMax : sto.ssc[psr[q]] / kSumMax;
        for (uint32_t a = 0; a < ((fq >> 8) & 0xFF); a++) {
          src_lds[(fq >> 16) + a] = off[q] + (uint64_t)a * np;
          if (Q) stp_lds[(fq >> 16) + a] = step;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < H; q++)
      if (lane + 64 * q < mcnt) {
        src[q] = src_lds[lane + 64 * q];
        if (Q) stp[q] = stp_lds[lane + 64 * q];
      }
  }
  // member m's record fields, read from the lane holding them (m is
  // wave-uniform)
#define LN32(x, m) (H == 1 || (m) < 64 ? lane_u32((x)[0], (m)) : lane_u32((x)[H - 1], (m) - 64))
#define LN64(x, m) (H == 1 || (m) < 64 ? lane_u64((x)[0], (m)) : lane_u64((x)[H - 1], (m) - 64))
#define LNF(x, m) (H == 1 || (m) < 64 ? lane_f32((x)[0], (m)) : lane_f32((x)[H - 1], (m) - 64))
  typedef MemberVec<CAP> SV;             // storage rows, by child; f32 member values
  typedef MemberVec<CAP, uint32_t> WV;   // f64 member values: low and high words
  typename SV::lo s0 = typename SV::lo(0.0f), l0 = typename SV::lo(0.0f);
  typename SV::mi s1 = typename SV::mi(0.0f), l1 = typename SV::mi(0.0f);
  typename SV::hi s2 = typename SV::hi(0.0f), l2 = typename SV::hi(0.0f);
  typename WV::lo w0 = typename WV::lo(0u), x0 = typename WV::lo(0u);
  typename WV::mi w1 = typename WV::mi(0u), x1 = typename WV::mi(0u);
  typename WV::hi w2 = typename WV::hi(0u), x2 = typename WV::hi(0u);
#define LV_GET(i) mval_get<V, CAP>(l0, l1, l2, w0, w1, w2, x0, x1, x2, (i))
#define LV_SET(i, x) mval_set<V, CAP>(l0, l1, l2, w0, w1, w2, x0, x1, x2, (i), (x))
#define SL_GET(i) SV::get(s0, s1, s2, (i))
#define SL_SET(i, x) SV::set(s0, s1, s2, (i), (x))
  const float* store = MODE == kCfr ? sto.r : sto.s;
  // per member (in its own lane): its value row with the chunk group's shift
  // applied, the top bit set unless the walk preloads it (a live terminal or
  // chance node in the span), so the unrolled loop below reads one word per
  // member
#pragma unroll
  for (int q = 0; q < H; q++)
    vslot[q] = B.v(vslot[q]) |
               ((lane + 64 * q < mcnt && (f[q] & 7) >= kChance && (f[q] & (kFLive | kFSpan)) == (kFLive | kFSpan))
                    ? 0u
                    : 0x80000000u);
#pragma unroll
  for (uint32_t m = 0; m < (uint32_t)CAP; m++) {  // unrolled: constant indices, all loads in flight
    if (m < mcnt) {
      const uint32_t vr = LN32(vslot, m);
      if (!(vr >> 31)) LV_SET(m, vals[(size_t)vr * rowlen + hh]);
      if (ROWS) {
        const uint64_t sm = LN64(src, m);
        if (sm != kNoRow) {
          if (Q) SL_SET(m, (MODE == kCfr ? (float)sto.rq[sm + hh] : (float)sto.sq[sm + hh]) * LNF(stp, m));
          else SL_SET(m, store[sm + hh]);
        }
      }
    }
  }
  float pos = 0.0f, neg = 0.0f;
  if (MODE == kCfr) {
    pos = prm->pos;
    neg = prm->neg;
  }
  for (uint32_t m = mcnt; m-- > 0;) {
    const uint32_t fm = LN32(f, m);
    if (!(fm & kFSpan)) continue;
    const uint32_t k = fm & 7, na = (fm >> 8) & 0xFF, cl = fm >> 16;
    if (qout != nullptr && (fm & kFRoot)) {
      // query: the values of the decision node's children (lib.rs:854-863)
      for (uint32_t a = 0; a < na; a++) {
        const V x = LV_GET(cl + a);
        if (act) qout[(size_t)a * np + h] = x;
      }
      return;
    }
    if (k >= kChance) continue;  // prefetched
    V v = V(0);
    if (fm & kFLive) {
      if ((int)k == p) {
        // passes over the actions re-read the registers (few values live at
        // once); the operations and their order are up_kernel's
        if constexpr (MODE == kCfr) {
          // one pass over the children for both sums (each accumulated in
          // action order, as up_kernel's separate loops do): the positive
          // regrets and the regret-weighted values
          float d = 0.0f, outv = 0.0f;
          for (uint32_t a = 0; a < na; a++) {
            const float r = pos0(SL_GET(cl + a));
            d += r;
            outv += r * LV_GET(cl + a);
          }
          const float inv = recip(d);
          outv *= inv;
          if (inv == 0.0f) {
            float sm = 0.0f;
            for (uint32_t a = 0; a < na; a++) sm += LV_GET(cl + a);
            outv = sm * (1.0f / (float)na);
          }
          // cfr.rs:374-379; the regret row of action a is the storage row of
          // child cl + a, at the node's offset + a * np (a macro, not a
          // lambda: the member vectors stay plain values)
#define UPDATED(a) (SL_GET(cl + (a)) * (SL_GET(cl + (a)) > 0.0f ? pos : neg) + LV_GET(cl + (a)) - outv)
          const uint64_t row0 = LN64(off, m);  // wave-uniform
          if (fm & kFFrozen) {
            // pinned regrets: no update (and, with 16-bit storage, no new scale)
          } else if (QP == 0) {
            if (act)
              for (uint32_t a = 0; a < na; a++) sto.r[row0 + (uint64_t)a * np + h] = UPDATED(a);
          } else if (QP == 1) {
            float lm = 0.0f;
            for (uint32_t a = 0; a < na; a++) {
              const float y = UPDATED(a);
              if (act) lm = fmaxf(lm, fabsf(y));
            }
            lm = wave_max(lm);
            if (lane == 0) atomicMax(&sto.rnx[LN32(psr, m)], __float_as_int(lm));
          } else {
            const float mx = __int_as_float(sto.rnx[LN32(psr, m)]);
            const float inv = mx > 0.0f ? kRegMax / mx : 0.0f;
            if (act)
              for (uint32_t a = 0; a < na; a++)
                sto.rq[row0 + (uint64_t)a * np + h] = mx > 0.0f ? enc_regret(UPDATED(a), inv) : (int16_t)0;
          }
#undef UPDATED
          v = outv;
        } else if constexpr (MODE == kBest) {
          // up_kernel's best response: the larger child value, the first on ties
          V b = LV_GET(cl);
          for (uint32_t a = 1; a < na; a++) {
            const V c = LV_GET(cl + a);
            b = b > c ? b : (c > b ? c : b);
          }
          v = b;
        } else {
          float d = SL_GET(cl);
          for (uint32_t a = 1; a < na; a++) d += SL_GET(cl + a);
          const float inv = recip(d), uni = 1.0f / (float)na;
          float acc = 0.0f;
          for (uint32_t a = 0; a < na; a++) acc += (inv == 0.0f ? uni : SL_GET(cl + a) * inv) * LV_GET(cl + a);
          v = acc;
        }
      } else {
        for (uint32_t a = 0; a < na; a++) v += LV_GET(cl + a);
      }
    }
    LV_SET(m, v);
  }
  const V top = LV_GET(0);
  if (act && QP != 2) {
    vals[(size_t)(LN32(vslot, 0u) & 0x7FFFFFFFu) * rowlen + h] = top;
    if (span.init && P.seg_pslot[(size_t)p * P.num_segs + seg] == NONE)  // a spot's root: its result row
      ((V*)span.res)[((size_t)P.mrec[moff].spot * 2 + p) * rowlen + h] = top;
  }
#undef LV_GET
#undef LV_SET
#undef SL_GET
#undef SL_SET
#undef LN32
#undef LN64
#undef LNF
}

template <typename V, int MODE, int CAP, int QP>
__global__ void __launch_bounds__(UP_T, CAP >= 64 ? 2 : CAP == 32 && QP != 2 ? 6 : (sizeof(V) == 8 && CAP == 48 ? 3 : 4)) up_reg_kernel(DPlan P, int p, uint32_t seg0, Span span, Bias B, V* __restrict__ vals,
                                                      uint32_t rowlen, const uint32_t* __restrict__ alive, Store sto,
                                                      const IterParams* __restrict__ prm, V* __restrict__ qout,
                                                      const uint32_t* __restrict__ list) {
  up_reg_walk<V, MODE, CAP, QP>(P, p, seg0, span, B, vals, rowlen, alive, sto, prm, qout, list);
}

// A range with many segments of 33 to 48 members: one launch over the list
// of segments of up to 48 members, each workgroup walking its segment with
// the 32-member code when it fits (one indexed move per member read) and the
// 48-member code otherwise (at the 48-member kernel's occupancy).
template <typename V, int MODE, int QP>
__global__ void __launch_bounds__(UP_T, 4) up_reg_dual_kernel(DPlan P, int p, uint32_t seg0, Span span, Bias B, V* __restrict__ vals,
                                                          uint32_t rowlen, const uint32_t* __restrict__ alive, Store sto,
                                                          const IterParams* __restrict__ prm, V* __restrict__ qout,
                                                          const uint32_t* __restrict__ list) {
  if (P.seg_head[list[seg0 + blockIdx.x]].y <= REG_T)
    up_reg_walk<V, MODE, (int)REG_T, QP>(P, p, seg0, span, B, vals, rowlen, alive, sto, prm, qout, list);
  else
    up_reg_walk<V, MODE, (int)REG_S, QP>(P, p, seg0, span, B, vals, rowlen, alive, sto, prm, qout, list);
}

// Values in LDS: the f64 average walk, and segments above REG_L members
// (list[seg0 + blockIdx.x] when a list is given, else seg0 + blockIdx.x).
// One wave per (segment, 64-hand tile), one hand per lane: values pulled up
// the segment's members in reverse breadth-first order (no barriers after
// staging: every lane owns its hand). Own nodes: cfr.rs:315-382 (regret
// update in CFR mode); opponent nodes: sum of the children, dead ones adding
// an exact zero (cfr.rs:421-432, river.rs:264-271); chance nodes: f64 sum
// over dealt cards, zero for own hands holding the card (cfr.rs:457-478).
// Every segment root's row is written (zeros when dead), so a chance node
// reads its children's rows unconditionally. QP as in up_reg_kernel.
template <typename V, int MODE, int QP>
__global__ void __launch_bounds__(UP_T) up_kernel(DPlan P, int p, uint32_t seg0, Span span, Bias B, V* __restrict__ vals,
                                                  uint32_t rowlen, const uint32_t* __restrict__ alive, Store sto,
                                                  const IterParams* __restrict__ prm, V* __restrict__ qout,
                                                  const uint32_t* __restrict__ list) {
  constexpr bool Q = QP != 0;
  extern __shared__ uint4 smem_up[];
  const uint32_t seg = list != nullptr ? list[seg0 + blockIdx.x] : seg0 + blockIdx.x;
  const uint32_t moff = P.seg_mem_off[seg], mcnt = P.seg_mem_cnt[seg];
  const SpotTab& S = P.spots[P.mrec[moff].spot];
  const uint32_t np = S.n[p], tid = threadIdx.x;
  const uint32_t h = blockIdx.y * UP_T + tid;
  const uint32_t root = P.mrec[moff].node;
  if (!span.has(root) && seg != span.seg) return;
  // LDS: the records, their alive flags, each decision member's value slot,
  // and one value per hand for each decision member (a terminal or chance
  // member is read from its row when its parent is walked, so the LDS per
  // wave scales with the decision members only)
  MemRec* rec = (MemRec*)smem_up;
  uint8_t* live = (uint8_t*)(rec + P.max_members);
  uint16_t* dslot = (uint16_t*)(live + ((P.max_members + 15) & ~15u));
  V* lv = (V*)(dslot + ((P.max_members + 7) & ~7u));
  stage_records(P.mrec + moff, mcnt, rec);
  __syncthreads();
  for (uint32_t m = tid; m < mcnt; m += UP_T) live[m] = alive[B.s(rec[m].slot[p])] != 0;
  {
    uint32_t base = 0;
    for (uint32_t m0 = 0; m0 < mcnt; m0 += UP_T) {
      const uint32_t m = m0 + tid;
      const bool dec = m < mcnt && rec[m].kind < kChance;
      const uint64_t bal = __ballot(dec);
      if (dec) dslot[m] = (uint16_t)(base + __popcll(bal & ((1ull << tid) - 1ull)));
      base += __popcll(bal);
    }
  }
  __syncthreads();
  // the max pass keeps every lane for its wave reductions: lanes past the
  // hands read hand 0 and store nothing
  const bool act = h < np;
  if (QP != 1 && !act) return;
  const uint32_t hc = act ? h : 0;
  // member c's value: a decision member's from LDS; a terminal or chance
  // member's from its row (0 when dead or outside the span)
  auto value = [&](uint32_t c) -> V {
    const MemRec& q = rec[c];
    if (q.kind >= kChance) return (live[c] && span.has(q.node)) ? vals[(size_t)B.v(q.vslot) * rowlen + hc] : V(0);
    return lv[(size_t)dslot[c] * UP_T + tid];
  };
  float pos = 0.0f, neg = 0.0f;
  if (MODE == kCfr) {
    pos = prm->pos;
    neg = prm->neg;
  }
  for (uint32_t m = mcnt; m-- > 0;) {
    const MemRec& r = rec[m];
    if (!span.has(r.node)) continue;
    const uint8_t k = r.kind;
    const uint32_t na = r.na, cl = r.clocal;
    if (QP != 1 && qout != nullptr && r.node == span.root) {
      // query: the values of the decision node's children (lib.rs:854-863)
      for (uint32_t a = 0; a < na; a++) qout[(size_t)a * np + h] = value(cl + a);
      return;
    }
    if (k >= kChance) continue;  // read by its parent
    V v = V(0);
    if (live[m]) {
      if ((int)k == p) {
        V cv[MAXA];
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) cv[a] = value(cl + a);
        const uint64_t off = r.offset;
        if (MODE == kCfr) {
          float rg[MAXA];
          const float rstep = Q ? sto.rsc[r.pslot] / kRegMax : 0.0f;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) rg[a] = Q ? (float)sto.rq[off + (size_t)a * np + hc] * rstep : sto.r[off + (size_t)a * np + hc];
          float d = pos0(rg[0]);
#pragma unroll
          for (int a = 1; a < MAXA; a++)
            if (a < (int)na) d += pos0(rg[a]);
          float outv = 0.0f;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) outv += pos0(rg[a]) * (float)cv[a];
          const float inv = recip(d);
          outv *= inv;
          if (inv == 0.0f) {
            float sm = 0.0f;
#pragma unroll
            for (int a = 0; a < MAXA; a++)
              if (a < (int)na) sm += (float)cv[a];
            outv = sm * (1.0f / (float)na);
          }
          float y[MAXA];
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) {
              const float x = rg[a];
              const float disc = x > 0.0f ? pos : neg;
              y[a] = x * disc + (float)cv[a] - outv;
            }
          if (r.tflags & kNodeFrozen) {
            // pinned regrets: no update (and, with 16-bit storage, no new scale)
          } else if (QP == 0) {
#pragma unroll
            for (int a = 0; a < MAXA; a++)
              if (a < (int)na) sto.r[off + (size_t)a * np + h] = y[a];
          } else if (QP == 1) {
            float lm = 0.0f;
#pragma unroll
            for (int a = 0; a < MAXA; a++)
              if (a < (int)na && act) lm = fmaxf(lm, fabsf(y[a]));
            lm = wave_max(lm);
            if (tid == 0) atomicMax(&sto.rnx[r.pslot], __float_as_int(lm));
          } else {
            const float mx = __int_as_float(sto.rnx[r.pslot]);
            const float inv = mx > 0.0f ? kRegMax / mx : 0.0f;
#pragma unroll
            for (int a = 0; a < MAXA; a++)
              if (a < (int)na) sto.rq[off + (size_t)a * np + h] = mx > 0.0f ? enc_regret(y[a], inv) : (int16_t)0;
          }
          v = (V)outv;
        } else if (MODE == kBest) {
          V b = cv[0];
#pragma unroll
          for (int a = 1; a < MAXA; a++)
            if (a < (int)na) b = b > cv[a] ? b : (cv[a] > b ? cv[a] : b);
          v = b;
        } else {
          V sv[MAXA];
          const float sstep = Q ? sto.ssc[r.pslot] / kSumMax : 0.0f;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) sv[a] = Q ? (V)((float)sto.sq[off + (size_t)a * np + h] * sstep) : (V)sto.s[off + (size_t)a * np + h];
          V d = sv[0];
#pragma unroll
          for (int a = 1; a < MAXA; a++)
            if (a < (int)na) d += sv[a];
          const V inv = recip(d), uni = V(1) / (V)na;
          V acc = V(0);
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) acc += (inv == V(0) ? uni : sv[a] * inv) * cv[a];
          v = acc;
        }
      } else {
        V cv[MAXA];
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) cv[a] = value(cl + a);
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) v += cv[a];
      }
    }
    lv[(size_t)dslot[m] * UP_T + tid] = v;
  }
  if (act && QP != 2) {
    const V v0 = value(0);
    vals[(size_t)B.v(rec[0].vslot) * rowlen + h] = v0;
    if (span.init && P.seg_pslot[(size_t)p * P.num_segs + seg] == NONE)  // a spot's root: its result row
      ((V*)span.res)[((size_t)rec[0].spot * 2 + p) * rowlen + h] = v0;
  }
}

// After a CFR traversal with 16-bit storage: every own node's new regret
// scale (its largest |new regret|) replaces the old; nodes the traversal
// left alone (dead) keep theirs (storage.rs:253-289 runs only for updated
// nodes).
__global__ void commit_scales_kernel(Store sto, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t b = sto.rnx[i];
  if (b >= 0) {
    sto.rsc[i] = __int_as_float(b);
    sto.rnx[i] = -1;
  }
}

// ------------------------------------------------------------------ leaves
// Depth-limited trees end the root street's betting in leaves whose value,
// per dealt card, comes from the value network. Per traversal of player p:
//   leaf_own_kernel    p's reach at every leaf (p's current strategy in CFR
//                      mode, the average strategy otherwise): one lane per
//                      hand walks the leaves' segment top-down;
//   leaf_input_kernel  one network input row per (leaf, dealt card): board
//                      mask, pot and stack fractions, both ranges at the leaf
//                      without the hands holding the card, each summing to 1;
//   (the network)      expected values per combo in pot fractions;
//   leaf_value_kernel  v(h) = norm * factor * pot * sum over cards c not in h
//                      of EV_p(c, h) * compat_c(h), compat_c(h) the
//                      opponent's reach over hands avoiding h and c
//                      (inclusion-exclusion over h's two cards and c), as the
//                      chance node the leaf replaces would sum its children.
template <typename V, bool AVG>
__global__ void __launch_bounds__(1024) leaf_own_kernel(DPlan P, int p, uint32_t seg, const float* __restrict__ regrets,
                                                        const float* __restrict__ sums, V* __restrict__ mreach,
                                                        const LeafRec* __restrict__ leaves, uint32_t nleaves,
                                                        V* __restrict__ own, uint32_t rowlen) {
  const uint32_t moff = P.seg_mem_off[seg], mcnt = P.seg_mem_cnt[seg];
  const SpotTab& S = P.spots[P.mrec[moff].spot];
  const uint32_t np = S.n[p];
  for (uint32_t h = threadIdx.x; h < np; h += blockDim.x) {
    mreach[h] = (V)S.weights[p][h];
    for (uint32_t m = 0; m < mcnt; m++) {
      const MemRec& r = P.mrec[moff + m];
      if (r.kind > kP1) continue;  // decision nodes only
      const V rr = mreach[(size_t)m * rowlen + h];
      const uint32_t na = r.na;
      V sg[MAXA];
      if ((int)r.kind == p) {
        if (AVG) {
          V sv[MAXA];
          V d = V(0);
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) {
              sv[a] = (V)sums[r.offset + (size_t)a * np + h];
              d += sv[a];
            }
          const V inv = recip(d), uni = V(1) / (V)na;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) sg[a] = inv == V(0) ? uni : sv[a] * inv;
        } else {
          float rg[MAXA];
          float d = 0.0f;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) {
              rg[a] = regrets[r.offset + (size_t)a * np + h];
              d += pos0(rg[a]);
            }
          const float inv = recip(d), uni = 1.0f / (float)na;
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) sg[a] = (V)(inv == 0.0f ? uni : pos0(rg[a]) * inv);
        }
      } else {
#pragma unroll
        for (int a = 0; a < MAXA; a++) sg[a] = V(1);
      }
#pragma unroll
      for (int a = 0; a < MAXA; a++)
        if (a < (int)na) mreach[(size_t)(r.clocal + a) * rowlen + h] = rr * sg[a];
    }
    for (uint32_t l = 0; l < nleaves; l++) own[(size_t)l * rowlen + h] = mreach[(size_t)leaves[l].local * rowlen + h];
  }
}

__device__ __forceinline__ uint32_t combo_of(uint32_t a, uint32_t b) {
  const uint32_t hi = a > b ? a : b, lo = a > b ? b : a;
  return hi * (hi - 1) / 2 + lo;
}

template <typename V>
__global__ void __launch_bounds__(256) leaf_input_kernel(DPlan P, int p, Bias B, const LeafRec* __restrict__ leaves,
                                                         uint32_t nleaves, const int8_t* __restrict__ deal, uint32_t ncards,
                                                         const int8_t* __restrict__ board, uint32_t board_len,
                                                         const V* __restrict__ own, const V* __restrict__ reach,
                                                         uint32_t rowlen, __bf16* __restrict__ rows) {
  __shared__ double red[256];
  const uint32_t r = blockIdx.x, tid = threadIdx.x;
  __bf16* row = rows + (size_t)r * 2752;
  for (uint32_t i = tid; i < 2752; i += 256) row[i] = (__bf16)0.0f;
  if (r >= nleaves * ncards) return;  // padding rows stay zero
  __syncthreads();
  const LeafRec L = leaves[r / ncards];
  const int c = deal[r % ncards];
  const SpotTab& S = P.spots[L.spot];
  if (tid < board_len) row[board[tid]] = (__bf16)1.0f;
  if (tid == 0) {
    row[c] = (__bf16)1.0f;
    const float tot = (float)L.pot + (float)L.stack;
    row[52] = (__bf16)((float)L.pot / tot);
    row[53] = (__bf16)((float)L.stack / tot);
  }
  for (int q = 0; q < 2; q++) {
    const V* rq = q == p ? own + (size_t)(r / ncards) * rowlen : reach + (size_t)B.s(L.slot[p]) * rowlen;
    const uint32_t nq = S.n[q];
    double s = 0.0;
    for (uint32_t h = tid; h < nq; h += 256)
      if (S.cards[q][2 * h] != c && S.cards[q][2 * h + 1] != c) s += (double)rq[h];
    red[tid] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
      if ((int)tid < k) red[tid] += red[tid + k];
      __syncthreads();
    }
    const double tot = red[0];
    __syncthreads();
    if (!(tot > 0.0)) continue;
    for (uint32_t h = tid; h < nq; h += 256) {
      const uint32_t a = S.cards[q][2 * h], b = S.cards[q][2 * h + 1];
      if ((int)a == c || (int)b == c) continue;
      row[54 + q * 1326 + combo_of(a, b)] = (__bf16)((double)rq[h] / tot);
    }
  }
}

template <typename V>
__global__ void __launch_bounds__(256) leaf_value_kernel(DPlan P, int p, Bias B, const LeafRec* __restrict__ leaves,
                                                         const int8_t* __restrict__ deal, uint32_t ncards,
                                                         const V* __restrict__ reach, const uint32_t* __restrict__ alive,
                                                         const float* __restrict__ ev, V* __restrict__ vals, uint32_t rowlen) {
  __shared__ double tot[52];
  __shared__ double red[256];
  const uint32_t li = blockIdx.x, tid = threadIdx.x;
  const LeafRec L = leaves[li];
  const SpotTab& S = P.spots[L.spot];
  const int o = p ^ 1;
  const uint32_t so = B.s(L.slot[p]);
  V* out = vals + (size_t)B.v(L.vslot) * rowlen;
  if (!alive[so]) return;  // dead: the row is never read
  const V* ro = reach + (size_t)so * rowlen;
  const uint32_t nq = S.n[o], np = S.n[p];
  double s = 0.0;
  for (uint32_t h = tid; h < nq; h += 256) s += (double)ro[h];
  red[tid] = s;
  if (tid < 52) {
    double t = 0.0;
    for (uint32_t k = S.bl_off[o][tid]; k < S.bl_off[o][tid + 1]; k++) t += (double)ro[S.bl_idx[o][k]];
    tot[tid] = t;
  }
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)tid < k) red[tid] += red[tid + k];
    __syncthreads();
  }
  const double total = red[0];
  const double scale = S.norm * (double)L.factor * (double)L.pot;
  for (uint32_t h = tid; h < np; h += 256) {
    const uint32_t c1 = S.cards[p][2 * h], c2 = S.cards[p][2 * h + 1];
    const uint32_t sm = S.same[p][h];
    const double same = sm == NONE ? 0.0 : (double)ro[sm];
    const uint32_t combo = combo_of(c1, c2);
    double v = 0.0;
    for (uint32_t ci = 0; ci < ncards; ci++) {
      const uint32_t c = (uint32_t)deal[ci];
      if (c == c1 || c == c2) continue;
      const uint32_t h1 = S.lookup[o][combo_of(c1, c)], h2 = S.lookup[o][combo_of(c2, c)];
      const double pair = (h1 == NONE ? 0.0 : (double)ro[h1]) + (h2 == NONE ? 0.0 : (double)ro[h2]);
      const double compat = total - tot[c1] - tot[c2] - tot[c] + same + pair;
      v += (double)ev[((size_t)li * ncards + ci) * 2652 + (size_t)p * 1326 + combo] * compat;
    }
    out[h] = (V)(v * scale);
  }
}

template <typename V>
__global__ void leaf_zero_kernel(Bias B, const LeafRec* __restrict__ leaves, V* __restrict__ vals, uint32_t rowlen) {
  V* out = vals + (size_t)B.v(leaves[blockIdx.x].vslot) * rowlen;
  for (uint32_t h = threadIdx.x; h < rowlen; h += blockDim.x) out[h] = V(0);
}

// ================================================================== engine
// Shape-specialised CFR walks (jit.hpp): per shape, its two kernels (one per
// traverser) and its segments, ascending (host copy; device copy).
struct JitLaunch {
  std::vector<hipFunction_t> fn;
  std::vector<std::vector<uint32_t>> lists;
  std::vector<const uint32_t*> dev;
  const float* zrow = nullptr;  // MAXA * rowlen zeros, read by dead members
};

struct Engine::DPlanHolder {
  DPlan d;
  std::vector<uint32_t> big;  // segments above REG_L members, ascending
  std::vector<uint32_t> m80;  // segments of REG_M + 1 to REG_L members, ascending
  std::vector<uint32_t> mid;  // segments of REG_S + 1 to REG_M members, ascending
  std::vector<uint32_t> m48;  // segments of REG_T + 1 to REG_S members, ascending
  std::vector<uint32_t> w32;  // segments with a walk, of up to REG_T members, ascending
  std::vector<uint32_t> w48;  // segments with a walk, of up to REG_S members, ascending
  // DCFR and best-response traversals with specialised walks: d with those
  // walks' segments taken out of the generic walk lists (cfr_* below), and
  // the walks
  DPlan dcfr;
  std::vector<uint32_t> cfr_m80, cfr_mid, cfr_m48, cfr_w32, cfr_w48;
  JitLaunch jit;
};

Engine::~Engine() {
  for (int i = 0; i < 3; i++)
    if (graph_[i]) (void)hipGraphExecDestroy(graph_[i]);
  for (auto& b : bufs_) (void)hipFree(b.p);
  if (params_host_) (void)hipHostFree(params_host_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

Status Engine::alloc(DeviceBuf& b, size_t bytes) {
  b.bytes = bytes;
  hipError_t e = hipMalloc(&b.p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    char m[200];
    std::snprintf(m, sizeof m, "out of GPU memory: needed %zu more bytes, %zu available", bytes, fr);
    set_error(m);
    set_error_detail(ErrorDetail{(int32_t)e, (uint64_t)bytes, (uint64_t)fr});
    return kOutOfMemory;
  }
  bufs_.push_back(b);
  bytes_ += bytes;
  return kOk;
}

template <typename T>
static size_t put(std::vector<unsigned char>& blob, const std::vector<T>& v) {
  size_t off = (blob.size() + 255) & ~size_t(255);
  blob.resize(off + v.size() * sizeof(T));
  if (!v.empty()) std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

Store Engine::store() const { return Store{regrets_, sums_, rq_, sq_, rsc_, ssc_, rnx_}; }

// Specialised walks (jit.hpp) for the frequent segment shapes: the DCFR walk
// (f32, or the two passes of 16-bit storage) and the FP64 best response. A shape is every
// member's kind (0 or 1 for a decision node, 2 for a member read from its
// row), action count and first child. A plan with at least kJitPlanSegments
// segments that have a walk compiles every shape of at least kJitShapeSegments
// of them (a smaller plan compiles none: its solve would not repay the
// compilation); PTS_JIT_MIN_SEGMENTS=n sets the per-shape count and drops the
// plan-size rule, PTS_JIT=0 turns them off. A spot's root segment, which also
// writes the spot's result row, keeps the generic walk. The other segments,
// the average-strategy walks and queries use the generic kernels.
constexpr uint32_t kJitPlanSegments = 8192, kJitShapeSegments = 64;
Status Engine::setup_jit(bool q16) {
  DPlanHolder& H = *dp_;
  const Plan& P = plan_;
  H.dcfr = H.d;
  const char* jo = std::getenv("PTS_JIT");
  if (jo && jo[0] == '0') return kOk;
  uint32_t min_segs = kJitShapeSegments;
  bool forced = false;
  if (const char* m = std::getenv("PTS_JIT_MIN_SEGMENTS")) {
    min_segs = (uint32_t)std::max(1, std::atoi(m));
    forced = true;
  }
  std::map<std::string, std::vector<uint32_t>> by_shape;
  std::map<std::string, jit::Shape> shape_of;
  size_t walk_segs = 0;
  for (uint32_t sg = 0; sg < (uint32_t)P.seg_mem_cnt.size(); sg++) {
    const uint32_t mc = P.seg_mem_cnt[sg];
    if (mc < 2 || mc > jit::kMaxMembers || P.seg_pslot[sg] == kNone) continue;
    const MemRec* rr = P.mrec.data() + P.seg_mem_off[sg];
    jit::Shape S(mc);
    std::string key(mc * 4, '\0');
    for (uint32_t m = 0; m < mc; m++) {
      const bool dec = is_player(rr[m].kind);
      S[m] = jit::Member{dec ? rr[m].kind : (uint8_t)2, dec ? rr[m].na : (uint8_t)0, dec ? rr[m].clocal : (uint16_t)0};
      std::memcpy(&key[m * 4], &S[m], 4);
    }
    auto& v = by_shape[key];
    if (v.empty()) shape_of[key] = S;
    v.push_back(sg);
    walk_segs++;
  }
  if (!forced && walk_segs < kJitPlanSegments) return kOk;
  std::vector<std::pair<size_t, std::string>> pick;
  for (const auto& kv : by_shape)
    if (kv.second.size() >= min_segs) pick.push_back({kv.second.size(), kv.first});
  std::sort(pick.begin(), pick.end(), [](const auto& x, const auto& y) { return x.first > y.first || (x.first == y.first && x.second < y.second); });
  if (pick.size() > 32) pick.resize(32);
  if (pick.empty()) return kOk;
  std::vector<jit::Shape> shapes;
  std::vector<uint8_t> taken(P.seg_mem_cnt.size(), 0);
  JitLaunch& J = H.jit;
  for (const auto& x : pick) {
    shapes.push_back(shape_of[x.second]);
    J.lists.push_back(std::move(by_shape[x.second]));
    for (uint32_t sg : J.lists.back()) taken[sg] = 1;
    jit_segments_ += J.lists.back().size();
  }
  Status st = jit::compile_walks(shapes, q16, J.fn, &jit_seconds_);
  if (st != kOk) {
    // the generic walks serve every segment, with the same results
    std::fprintf(stderr, "[pts] specialised walks unavailable, the generic walks run instead: %s\n", last_error().c_str());
    J = JitLaunch();
    jit_segments_ = 0;
    jit_seconds_ = 0;
    set_error("");
    return kOk;
  }
  auto rest = [&](const std::vector<uint32_t>& l, std::vector<uint32_t>& out) {
    for (uint32_t sg : l)
      if (!taken[sg]) out.push_back(sg);
  };
  rest(H.m80, H.cfr_m80);
  rest(H.mid, H.cfr_mid);
  rest(H.m48, H.cfr_m48);
  rest(H.w32, H.cfr_w32);
  rest(H.w48, H.cfr_w48);
  DeviceBuf z;
  if ((st = alloc(z, (size_t)MAXA * rowlen_ * sizeof(float))) != kOk) return st;
  HIPCHK(hipMemset(z.p, 0, z.bytes));
  J.zrow = (const float*)z.p;
  // one buffer: the shapes' lists, then the reduced generic lists
  std::vector<const std::vector<uint32_t>*> all;
  for (const auto& l : J.lists) all.push_back(&l);
  all.push_back(&H.cfr_m80);
  all.push_back(&H.cfr_mid);
  all.push_back(&H.cfr_m48);
  all.push_back(&H.cfr_w32);
  all.push_back(&H.cfr_w48);
  size_t n = 0;
  for (const auto* l : all) n += l->size() + 64;  // each list on its own 256-byte boundary
  DeviceBuf b;
  if ((st = alloc(b, n * 4)) != kOk) return st;
  std::vector<uint32_t> host(n, 0);
  std::vector<const uint32_t*> dev;
  size_t at = 0;
  for (const auto* l : all) {
    std::copy(l->begin(), l->end(), host.begin() + at);
    dev.push_back((const uint32_t*)b.p + at);
    at += (l->size() + 63) & ~size_t(63);
  }
  HIPCHK(hipMemcpy(b.p, host.data(), n * 4, hipMemcpyHostToDevice));
  const size_t k = J.lists.size();
  J.dev.assign(dev.begin(), dev.begin() + k);
  DPlan& c = H.dcfr;
  c.m80 = dev[k + 0];
  c.mid = dev[k + 1];
  c.m48 = dev[k + 2];
  c.w32 = dev[k + 3];
  c.w48 = dev[k + 4];
  c.m80_host = H.cfr_m80.data();
  c.mid_host = H.cfr_mid.data();
  c.m48_host = H.cfr_m48.data();
  c.w32_host = H.cfr_w32.data();
  c.w48_host = H.cfr_w48.data();
  c.num_m80 = (uint32_t)H.cfr_m80.size();
  c.num_mid = (uint32_t)H.cfr_mid.size();
  c.num_m48 = (uint32_t)H.cfr_m48.size();
  c.num_w32 = (uint32_t)H.cfr_w32.size();
  c.num_w48 = (uint32_t)H.cfr_w48.size();
  jit_shapes_ = (uint32_t)k;
  if (const char* vb = std::getenv("PTS_JIT_VERBOSE"))
    if (vb[0] == '1')
      std::fprintf(stderr, "[pts] specialised walks: %u shapes, %llu segments, compiled in %.2f s\n", jit_shapes_,
                   (unsigned long long)jit_segments_, jit_seconds_);
  return kOk;
}

Status Engine::create(int device, Plan&& plan, const float* regrets, const float* sums, int precision,
                      uint64_t scratch_budget, std::unique_ptr<Engine>& out, bool q16) {
  std::unique_ptr<Engine> e(new Engine());
  if (precision != kFast && precision != kExact) {
    set_error("invalid input: precision");
    return kInvalidInput;
  }
  if (q16 && (plan.num_pslots == 0 || regrets || sums)) {
    set_error("invalid input: 16-bit storage needs player slots in the plan and starts from zeros");
    return kInvalidInput;
  }
  e->q16_ = q16;
  e->precision_ = precision;
  e->plan_ = std::move(plan);
  e->device_ = device;
  Plan& P = e->plan_;
  for (size_t i = 0; i < P.num_nodes; i++)
    if (is_player(P.kind[i]) && P.na[i] > MAXA) {
      set_error("invalid input: a decision node has more than 8 actions");
      return kInvalidInput;
    }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    set_error("no GPU device");
    return kNoDevice;
  }
  if (device >= count) {
    set_error("no such GPU device");
    return kNoDevice;
  }
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipStreamCreateWithFlags(&e->stream_, hipStreamNonBlocking));
  // the records are copied into LDS as 16-byte words
  static_assert(sizeof(MemRec) % 16 == 0, "MemRec staging");

  // value rows start on 128-byte boundaries
  const uint32_t rowlen = (std::max(P.n[0], P.n[1]) + 31u) & ~31u;
  e->rowlen_ = rowlen;
  // the storage must fit with a 512 MiB margin (crates/solver/src/gpu/engine.rs:261-315)
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  const size_t margin = 512ull << 20;
  // f32 storage: two 4-byte cells per storage cell; 16-bit storage: two
  // 2-byte cells plus three 4-byte words per decision node (its two scales
  // and the regrets' next scale), as the allocation below
  const size_t base_need = 2 * P.storage_len * (q16 ? 2 : sizeof(float)) + (q16 ? (size_t)P.num_pslots * 12 : 0) +
                           P.num_nodes * 16 + (size_t)P.num_boards * (P.n[0] + P.n[1]) * 24;
  if (base_need + margin > fr) {
    const size_t avail = fr > margin ? fr - margin : 0;
    char m[200];
    std::snprintf(m, sizeof m, "out of GPU memory: the storage needs %zu bytes, %zu available", base_need, avail);
    set_error(m);
    set_error_detail(ErrorDetail{0, (uint64_t)base_need, (uint64_t)avail});
    return kOutOfMemory;
  }

  // plan blob: the records, then every spot's hands and strength tables
  std::vector<unsigned char> blob;
  size_t o_card = put(blob, P.card), o_smo = put(blob, P.seg_mem_off), o_smc = put(blob, P.seg_mem_cnt);
  size_t o_sps = put(blob, P.seg_pslot), o_sf = put(blob, P.seg_factor), o_mr = put(blob, P.mrec), o_tr = put(blob, P.trec);
  size_t o_cr = put(blob, P.crec), o_ccv = put(blob, P.chance_child_vslot);
  size_t o_cio = put(blob, P.ciso_off), o_cis = put(blob, P.ciso);
  // room for the most runs (runs of one terminal), so the run length can be retuned
  std::vector<TermBatch> tb_room(std::max<size_t>(P.trec.size(), P.tbatch.size()) + 1);
  std::copy(P.tbatch.begin(), P.tbatch.end(), tb_room.begin());
  size_t o_tb = put(blob, tb_room);
  const uint32_t K = P.num_spots;
  struct SpotOff {
    size_t cards[2], same[2], bo[2], bi[2], eo[2], ec[2], cb[2], ei[2], co[2], ci[2], hc[2], w[2], lk[2], pm[2];
  };
  std::vector<SpotOff> so(K);
  for (uint32_t k = 0; k < K; k++) {
    const Plan& T = P.spot(k);
    for (int q = 0; q < 2; q++) {
      so[k].cards[q] = put(blob, T.cards[q]);
      so[k].same[q] = put(blob, T.same[q]);
      so[k].bo[q] = put(blob, T.bl_off[q]);
      so[k].bi[q] = put(blob, T.bl_idx[q]);
      so[k].eo[q] = put(blob, T.ent_off[q]);
      so[k].ec[q] = put(blob, T.ent_cnt[q]);
      so[k].cb[q] = put(blob, T.cl_base[q]);
      so[k].ei[q] = put(blob, T.ent_idx[q]);
      so[k].co[q] = put(blob, T.cl_off[q]);
      so[k].ci[q] = put(blob, T.cl_idx[q]);
      so[k].hc[q] = put(blob, T.hand_cnt[q]);
      so[k].w[q] = put(blob, T.weights[q]);
      so[k].lk[q] = put(blob, T.lookup[q]);
      so[k].pm[q] = put(blob, T.iso_perm[q]);
    }
  }
  std::vector<uint32_t> ones(K, 1u);
  const size_t o_act = put(blob, ones);
  // down-pass groups: per segment and traverser, the opponent's decision
  // nodes by level (the number of the opponent's decision nodes above them
  // in the segment: a level's nodes read only rows of earlier levels), cut
  // into groups of at most DG_N nodes and DG_A actions (DG_A16 with 16-bit storage)
  std::vector<uint32_t> dg_off[2], dg_ent[2];
  for (int q = 0; q < 2; q++) {
    const int opp = q ^ 1;
    dg_off[q].assign(P.seg_mem_cnt.size() + 1, 0);
    std::vector<uint32_t> lvl;
    std::vector<std::pair<uint32_t, uint32_t>> ent;  // (level, local)
    for (uint32_t sg = 0; sg < (uint32_t)P.seg_mem_cnt.size(); sg++) {
      dg_off[q][sg] = (uint32_t)dg_ent[q].size();
      const MemRec* rr = P.mrec.data() + P.seg_mem_off[sg];
      const uint32_t mc = P.seg_mem_cnt[sg];
      lvl.assign(mc, 0);
      ent.clear();
      for (uint32_t m = 0; m < mc; m++) {  // members in breadth-first order: parents first
        if (!is_player(rr[m].kind)) continue;
        const uint32_t up = lvl[m] + ((int)rr[m].kind == opp ? 1u : 0u);
        for (uint32_t a = 0; a < rr[m].na; a++) lvl[rr[m].clocal + a] = up;
        if ((int)rr[m].kind == opp) ent.push_back({lvl[m], m});
      }
      std::stable_sort(ent.begin(), ent.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      uint32_t grp = 0, acts = 0, nodes = 0, cur = ~0u;
      for (const auto& x : ent) {
        const uint32_t na = rr[x.second].na;
        if (nodes > 0 && (x.first != cur || acts + na > (uint32_t)(q16 ? DG_A16 : DG_A) || nodes == (uint32_t)DG_N)) {
          grp++;
          acts = nodes = 0;
        }
        cur = x.first;
        acts += na;
        nodes++;
        dg_ent[q].push_back(x.second | grp << 16);
      }
    }
    dg_off[q][P.seg_mem_cnt.size()] = (uint32_t)dg_ent[q].size();
  }
  const size_t o_dgo0 = put(blob, dg_off[0]), o_dgo1 = put(blob, dg_off[1]);
  const size_t o_dge0 = put(blob, dg_ent[0]), o_dge1 = put(blob, dg_ent[1]);
  std::vector<uint32_t> big, m80, mid, m48, w32, w48;
  std::vector<uint4> head(P.seg_mem_cnt.size());
  for (uint32_t sg = 0; sg < (uint32_t)P.seg_mem_cnt.size(); sg++) {
    const MemRec& r0 = P.mrec[P.seg_mem_off[sg]];
    // a chance child (it has a parent reach slot) of one terminal or chance
    // node: that member, the segment root, writes the root row itself
    const bool walk = !(P.seg_mem_cnt[sg] == 1 && P.seg_pslot[sg] != kNone && (r0.kind == kChance || is_terminal(r0.kind)));
    if (P.seg_mem_cnt[sg] > REG_L) big.push_back(sg);
    else if (P.seg_mem_cnt[sg] > REG_M) m80.push_back(sg);
    else if (P.seg_mem_cnt[sg] > REG_S) mid.push_back(sg);
    else if (P.seg_mem_cnt[sg] > REG_T) m48.push_back(sg);
    if (walk && P.seg_mem_cnt[sg] <= REG_T) w32.push_back(sg);
    if (walk && P.seg_mem_cnt[sg] <= REG_S) w48.push_back(sg);
    const Plan& T = P.spot(r0.spot);
    head[sg] = uint4{P.seg_mem_off[sg], P.seg_mem_cnt[sg], r0.node, T.n[0] | T.n[1] << 16};
  }
  const size_t o_big = put(blob, big), o_m80 = put(blob, m80), o_mid = put(blob, mid), o_m48 = put(blob, m48),
               o_w32 = put(blob, w32), o_w48 = put(blob, w48), o_head = put(blob, head);
  const size_t o_spots = (blob.size() + 255) & ~size_t(255);
  blob.resize(o_spots + K * sizeof(SpotTab));
  DeviceBuf b;
  Status st;
  if ((st = e->alloc(b, blob.size())) != kOk) return st;
  unsigned char* base = (unsigned char*)b.p;
#define AT(T, o) ((const T*)(base + (o)))
  std::vector<SpotTab> tabs(K);
  for (uint32_t k = 0; k < K; k++) {
    const Plan& T = P.spot(k);
    SpotTab& t = tabs[k];
    t.n[0] = T.n[0];
    t.n[1] = T.n[1];
    t.root_slot[0] = P.spot_root_slot[0][k];
    t.root_slot[1] = P.spot_root_slot[1][k];
    t.root_vslot = P.spot_root_vslot[k];
    t.pad = 0;
    t.half_pot = T.half_pot;
    t.norm = T.norm;
    for (int q = 0; q < 2; q++) {
      t.weights[q] = AT(float, so[k].w[q]);
      t.cards[q] = AT(uint8_t, so[k].cards[q]); t.same[q] = AT(uint32_t, so[k].same[q]);
      t.bl_off[q] = AT(uint32_t, so[k].bo[q]); t.bl_idx[q] = AT(uint16_t, so[k].bi[q]);
      t.ent_off[q] = AT(uint32_t, so[k].eo[q]); t.ent_cnt[q] = AT(uint32_t, so[k].ec[q]);
      t.cl_base[q] = AT(uint32_t, so[k].cb[q]); t.ent_idx[q] = AT(uint16_t, so[k].ei[q]);
      t.cl_off[q] = AT(uint16_t, so[k].co[q]); t.cl_idx[q] = AT(uint16_t, so[k].ci[q]);
      t.hand_cnt[q] = AT(uint4, so[k].hc[q]);
      t.lookup[q] = AT(uint32_t, so[k].lk[q]);
      t.perm[q] = T.iso_perm[q].empty() ? nullptr : AT(uint16_t, so[k].pm[q]);
    }
  }
  std::memcpy(blob.data() + o_spots, tabs.data(), K * sizeof(SpotTab));
  HIPCHK(hipMemcpy(b.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
  e->dp_.reset(new DPlanHolder());
  DPlan& d = e->dp_->d;
  d.num_nodes = (uint32_t)P.num_nodes;
  d.num_segs = (uint32_t)P.seg_root.size();
  d.num_spots = K;
  d.max_entries = std::max(P.max_entries, 1u);
  d.max_members = std::max(P.max_members, 1u);
  {
    uint32_t md = 1;
    for (size_t g = 0; g < P.seg_mem_cnt.size(); g++) {
      uint32_t c = 0;
      for (uint32_t m = 0; m < P.seg_mem_cnt[g]; m++) c += is_player(P.mrec[P.seg_mem_off[g] + m].kind) ? 1u : 0u;
      md = std::max(md, c);
    }
    d.max_dec = md;
  }
  d.card = AT(int8_t, o_card);
  d.seg_mem_off = AT(uint32_t, o_smo); d.seg_mem_cnt = AT(uint32_t, o_smc); d.seg_pslot = AT(uint32_t, o_sps);
  d.seg_factor = AT(float, o_sf); d.mrec = AT(MemRec, o_mr); d.trec = AT(TermRec, o_tr);
  d.crec = AT(MemRec, o_cr); d.chance_child_vslot = AT(uint32_t, o_ccv); d.tbatch = AT(TermBatch, o_tb);
  d.ciso_off = AT(uint32_t, o_cio); d.ciso = AT(IsoRef, o_cis);
  d.spots = AT(SpotTab, o_spots);
  e->tbatch_dev_ = (TermBatch*)(base + o_tb);
  e->run_length_ = default_run_length(P);
  // down workgroups: when a depth has many segments (the flop's river level
  // has 670,320), workgroups of 128 threads keep more segments in flight per
  // CU (flop down pass 67 -> 55 ms per iteration against 512); a few hundred
  // segments (a turn's river level) run best at 512 (the turn: 0.38 against
  // 0.48 ms per iteration at 128). 16-bit storage keeps 512: its down pass
  // stages a group's new sums in LDS (up to 16 rows per workgroup), so small
  // workgroups would leave a CU a few hundred threads (the flop's 16-bit down
  // pass: 392 ms at 128 against 131 at 512)
  {
    uint32_t widest = 0;
    for (size_t dd = 0; dd + 1 < P.depth_off.size(); dd++) widest = std::max(widest, P.depth_off[dd + 1] - P.depth_off[dd]);
    e->down_threads_ = widest >= 16384 && !q16 ? 128u : 512u;
  }
  d.active = AT(uint32_t, o_act);
  e->active_ = (uint32_t*)(base + o_act);
  e->dp_->big = std::move(big);
  e->dp_->m80 = std::move(m80);
  e->dp_->mid = std::move(mid);
  e->dp_->m48 = std::move(m48);
  e->dp_->w32 = std::move(w32);
  e->dp_->w48 = std::move(w48);
  d.w32 = AT(uint32_t, o_w32);
  d.w48 = AT(uint32_t, o_w48);
  d.w32_host = e->dp_->w32.data();
  d.num_w32 = (uint32_t)e->dp_->w32.size();
  d.w48_host = e->dp_->w48.data();
  d.num_w48 = (uint32_t)e->dp_->w48.size();
  d.big = AT(uint32_t, o_big);
  d.m80 = AT(uint32_t, o_m80);
  d.mid = AT(uint32_t, o_mid);
  d.m48 = AT(uint32_t, o_m48);
  d.seg_head = AT(uint4, o_head);
  d.dg_off[0] = AT(uint32_t, o_dgo0);
  d.dg_off[1] = AT(uint32_t, o_dgo1);
  d.dg_ent[0] = AT(uint32_t, o_dge0);
  d.dg_ent[1] = AT(uint32_t, o_dge1);
  d.big_host = e->dp_->big.data();
  d.num_big = (uint32_t)e->dp_->big.size();
  d.m80_host = e->dp_->m80.data();
  d.num_m80 = (uint32_t)e->dp_->m80.size();
  d.mid_host = e->dp_->mid.data();
  d.num_mid = (uint32_t)e->dp_->mid.size();
  d.m48_host = e->dp_->m48.data();
  d.num_m48 = (uint32_t)e->dp_->m48.size();
#undef AT
  // storage, then the scratch rows: chunk groups sized to what is left
  DeviceBuf br, bs, bre, bp, bm, bsc;
  const size_t cell = q16 ? 2 : sizeof(float);
  if ((st = e->alloc(br, P.storage_len * cell)) != kOk) return st;
  if ((st = e->alloc(bs, P.storage_len * cell)) != kOk) return st;
  if (q16) {
    if ((st = e->alloc(bsc, (size_t)P.num_pslots * 12)) != kOk) return st;
    e->rq_ = (int16_t*)br.p;
    e->sq_ = (uint16_t*)bs.p;
    e->rsc_ = (float*)bsc.p;
    e->ssc_ = e->rsc_ + P.num_pslots;
    e->rnx_ = (int32_t*)(e->ssc_ + P.num_pslots);
  } else {
    e->regrets_ = (float*)br.p;
    e->sums_ = (float*)bs.p;
  }
  if ((st = e->alloc(bm, P.mrec.size() + 256)) != kOk) return st;
  d.malive = (uint8_t*)bm.p;
  if ((st = e->alloc(bp, sizeof(IterParams) + (size_t)K * 2 * rowlen * 8)) != kOk) return st;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  uint64_t budget = fr > margin ? fr - margin : 0;
  if (scratch_budget != 0) budget = std::min<uint64_t>(budget, scratch_budget);
  e->budget_ = budget;
  // One scratch pool; the f32 (DCFR) and f64 (best response, average) row
  // types each lay out their reach rows, value rows and alive flags in it
  // (they never run at the same time), grouped to fit the budget.
  size_t lay[2][3];
  for (int f64 = 0; f64 < 2; f64++) {
    ScratchRows r = partition_chunks(P, f64 ? 8 : 4, rowlen, budget, e->groups_[f64]);
    std::string err = check_groups(P, e->groups_[f64], r);
    if (!err.empty()) {
      set_error("invalid input: " + err);
      return kInvalidInput;
    }
    lay[f64][0] = 0;
    lay[f64][1] = (r.reach_bytes + 255) & ~size_t(255);
    lay[f64][2] = lay[f64][1] + ((r.val_bytes + 255) & ~size_t(255));
    e->pool_bytes_[f64] = lay[f64][2] + r.alive_rows * 4;
  }
  const size_t pool = std::max(e->pool_bytes_[0], e->pool_bytes_[1]);
  if ((st = e->alloc(bre, pool)) != kOk) return st;
  for (int f64 = 0; f64 < 2; f64++) {
    e->reach_[f64] = (char*)bre.p + lay[f64][0];
    e->vals_[f64] = (char*)bre.p + lay[f64][1];
    e->alive_[f64] = (uint32_t*)((char*)bre.p + lay[f64][2]);
  }
  e->params_ = (IterParams*)bp.p;
  HIPCHK(hipHostMalloc((void**)&e->params_host_, sizeof(IterParams)));
  if ((st = e->upload(regrets, sums)) != kOk) return st;

  // LDS sizes
  const size_t mm = std::max<size_t>(P.max_members, 1);
  e->lds_down_ = mm * sizeof(MemRec) + ((mm + 15) & ~size_t(15)) + 4 * mm + (q16 ? (size_t)DG_A16 * rowlen * 4 : 0);
  for (int f64 = 0; f64 < 2; f64++) {
    size_t vs = f64 ? 8 : 4;
    // the fast terminal kernel's LDS is static (TermLds)
    e->lds_term_[f64] = precision == kExact ? (size_t)rowlen * 3 * 8 + (size_t)rowlen * vs : 0;
    // records, alive flags, value slots, and one value row per decision member
    e->lds_up_[f64] = mm * sizeof(MemRec) + ((mm + 15) & ~size_t(15)) + ((mm + 7) & ~size_t(7)) * 2 +
                      (size_t)d.max_dec * UP_T * vs;
  }
  size_t lds_max = std::max({e->lds_term_[1], e->lds_up_[1], e->lds_down_});
  if (lds_max > 159 * 1024) {  // 160 KiB per workgroup, less the kernels' static arrays
    set_error("invalid input: a segment or strength table is too large for the LDS");
    return kInvalidInput;
  }
  // dynamic LDS beyond the default 64 KiB is requested per kernel, for
  // exactly the bytes it is launched with
  auto lds_attr = [&](const void* f, size_t bytes) -> Status {
    if (bytes > 64 * 1024) HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return kOk;
  };
  const bool fastp = precision == kFast;
  if ((st = lds_attr(fastp ? (const void*)terminal_fast_kernel<double, 0> : (const void*)terminal_kernel<double>,
                     e->lds_term_[1])) != kOk ||
      (st = lds_attr(fastp ? (const void*)terminal_fast_kernel<double, 1> : (const void*)terminal_kernel<double>,
                     e->lds_term_[1])) != kOk ||
      (st = lds_attr(fastp ? (const void*)terminal_fast_kernel<float, 0> : (const void*)terminal_kernel<float>,
                     e->lds_term_[0])) != kOk ||
      (st = lds_attr(fastp ? (const void*)terminal_fast_kernel<float, 1> : (const void*)terminal_kernel<float>,
                     e->lds_term_[0])) != kOk ||
      (st = lds_attr((const void*)up_kernel<double, kBest, 0>, e->lds_up_[1])) != kOk ||
      (st = lds_attr((const void*)up_kernel<double, kAverage, 0>, e->lds_up_[1])) != kOk ||
      (st = lds_attr((const void*)up_kernel<float, kAverage, 0>, e->lds_up_[0])) != kOk ||
      (st = lds_attr((const void*)up_kernel<float, kCfr, 0>, e->lds_up_[0])) != kOk ||
      (st = lds_attr((const void*)down_kernel<float, kCfr, false>, e->lds_down_)) != kOk ||
      (st = lds_attr((const void*)down_kernel<double, kAverage, false>, e->lds_down_)) != kOk ||
      (st = lds_attr((const void*)down_kernel<float, kAverage, false>, e->lds_down_)) != kOk)
    return st;
  if (q16 && ((st = lds_attr((const void*)up_kernel<double, kBest, 3>, e->lds_up_[1])) != kOk ||
              (st = lds_attr((const void*)up_kernel<double, kAverage, 3>, e->lds_up_[1])) != kOk ||
              (st = lds_attr((const void*)up_kernel<float, kAverage, 3>, e->lds_up_[0])) != kOk ||
              (st = lds_attr((const void*)up_kernel<float, kCfr, 1>, e->lds_up_[0])) != kOk ||
              (st = lds_attr((const void*)up_kernel<float, kCfr, 2>, e->lds_up_[0])) != kOk ||
              (st = lds_attr((const void*)down_kernel<float, kCfr, true>, e->lds_down_)) != kOk ||
              (st = lds_attr((const void*)down_kernel<double, kAverage, true>, e->lds_down_)) != kOk ||
              (st = lds_attr((const void*)down_kernel<float, kAverage, true>, e->lds_down_)) != kOk))
    return st;
  // the specialised walks take a copy of the finished device view (d)
  if ((st = e->setup_jit(q16)) != kOk) return st;
  const char* ng = std::getenv("PTS_NO_GRAPH");
  e->use_graph_ = !(ng && ng[0] == '1');
  if (e->use_graph_ && (st = e->build_graphs()) != kOk) return st;
  out = std::move(e);
  return kOk;
}

Status Engine::upload(const float* regrets, const float* sums) {
  if (q16_) {
    if (regrets || sums) {
      set_error("invalid input: 16-bit storage takes its integers and scales (upload_q)");
      return kInvalidInput;
    }
    const size_t bytes = plan_.storage_len * 2;
    HIPCHK(hipMemset(rq_, 0, bytes));
    HIPCHK(hipMemset(sq_, 0, bytes));
    HIPCHK(hipMemset(rsc_, 0, (size_t)plan_.num_pslots * 8));
    HIPCHK(hipMemset(rnx_, 0xFF, (size_t)plan_.num_pslots * 4));  // -1: no new scale
    return kOk;
  }
  const size_t bytes = plan_.storage_len * sizeof(float);
  if (regrets) HIPCHK(hipMemcpy(regrets_, regrets, bytes, hipMemcpyHostToDevice));
  else HIPCHK(hipMemset(regrets_, 0, bytes));
  if (sums) HIPCHK(hipMemcpy(sums_, sums, bytes, hipMemcpyHostToDevice));
  else HIPCHK(hipMemset(sums_, 0, bytes));
  return kOk;
}


Status Engine::download_q(int16_t* regrets, uint16_t* sums, float* regret_scale, float* strat_scale) const {
  if (!q16_) {
    set_error("invalid input: f32 storage");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  if (regrets) HIPCHK(hipMemcpy(regrets, rq_, plan_.storage_len * 2, hipMemcpyDeviceToHost));
  if (sums) HIPCHK(hipMemcpy(sums, sq_, plan_.storage_len * 2, hipMemcpyDeviceToHost));
  if (regret_scale) HIPCHK(hipMemcpy(regret_scale, rsc_, (size_t)plan_.num_pslots * 4, hipMemcpyDeviceToHost));
  if (strat_scale) HIPCHK(hipMemcpy(strat_scale, ssc_, (size_t)plan_.num_pslots * 4, hipMemcpyDeviceToHost));
  return kOk;
}

Status Engine::upload_q(const int16_t* regrets, const uint16_t* sums, const float* regret_scale, const float* strat_scale) {
  if (!q16_ || !regrets || !sums || !regret_scale || !strat_scale) {
    set_error("invalid input: 16-bit storage arrays");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  HIPCHK(hipMemcpy(rq_, regrets, plan_.storage_len * 2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(sq_, sums, plan_.storage_len * 2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(rsc_, regret_scale, (size_t)plan_.num_pslots * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ssc_, strat_scale, (size_t)plan_.num_pslots * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(rnx_, 0xFF, (size_t)plan_.num_pslots * 4));
  return kOk;
}

Status Engine::download(float* regrets, float* sums) const {
  if (q16_) {
    set_error("invalid input: 16-bit storage (download_q)");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  HIPCHK(hipMemcpy(regrets, regrets_, plan_.storage_len * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sums, sums_, plan_.storage_len * sizeof(float), hipMemcpyDeviceToHost));
  return kOk;
}

Status Engine::read_sums(uint64_t offset, size_t len, float* out, uint32_t pslot) const {
  if (offset + len > plan_.storage_len || (q16_ && pslot >= plan_.num_pslots)) {
    set_error("invalid input: storage range");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  if (q16_) {  // storage.rs:236-246
    std::vector<uint16_t> q(len);
    float scale = 0.0f;
    HIPCHK(hipMemcpy(q.data(), sq_ + offset, len * 2, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(&scale, ssc_ + pslot, 4, hipMemcpyDeviceToHost));
    const float step = scale / 65535.0f;
    for (size_t i = 0; i < len; i++) out[i] = (float)q[i] * step;
    return kOk;
  }
  HIPCHK(hipMemcpy(out, sums_ + offset, len * sizeof(float), hipMemcpyDeviceToHost));
  return kOk;
}

// The root-value row of traverser p is kept past the other traversal.
// Root value rows of every spot, kept past the traversals: (spot, p) at
// byte offset sizeof(IterParams) + (spot * 2 + p) * rowlen * element size.
static inline size_t result_offset(uint32_t spot, int p, uint32_t rowlen, size_t vs) {
  return sizeof(IterParams) + ((size_t)spot * 2 + p) * rowlen * vs;
}

struct ProfRec {
  hipStream_t st;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  void start(int cls) {
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, st);
    ev.push_back({cls, {a, b}});
  }
  void stop() { (void)hipEventRecord(ev.back().second.second, st); }
  ~ProfRec() {
    for (auto& e : ev) {
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
  }
};
#define PROF_START(c) \
  if (prof) prof->start(c)
#define PROF_STOP() \
  if (prof) prof->stop()

// Q: 16-bit storage (a CFR up pass then walks twice, QP 1 and 2; reads use QP 3)
template <typename V, int DMODE, int UMODE, bool Q>
static void launch_span(const DPlan& d, const Plan& P, hipStream_t st, int p, Span span, uint32_t first_depth,
                        const std::vector<ChunkGroup>& groups, void* reach, uint32_t* alive, const Store& sto,
                        const IterParams* prm, void* vals, uint32_t rowlen, size_t lds_down, size_t lds_term, size_t lds_up,
                        V* qout, bool fast, uint32_t down_t, ProfRec* prof = nullptr,
                        const std::function<void()>* leaf_hook = nullptr, const JitLaunch* jit = nullptr) {
  const uint32_t D = P.depths(), C = P.num_chunks;
  const uint32_t tiles = (P.n[p] + UP_T - 1) / UP_T;
  const Bias B0{P.slot_s0[p], 0, P.val_v0, 0};
  auto down = [&](uint32_t a, uint32_t b, const Bias& B) {
    if (b <= a) return;
    // a launch of few segments (the top of the tree, a batch's 320 spot
    // roots) cannot fill the chip with workgroups: 1024 threads per segment
    // (the turn's 721 river segments run slower so: 0.135 against 0.115 ms)
    const uint32_t threads = (b - a) < 512 ? 1024 : down_t;
    PROF_START(kKDown);
    hipLaunchKernelGGL((down_kernel<V, DMODE, Q>), dim3(b - a), dim3(threads), lds_down, st, d, p, a, span, B, (V*)reach, alive,
                       sto, prm, rowlen);
    PROF_STOP();
  };
  auto terminals = [&](uint32_t a, uint32_t b, const Bias& B) {
    if (b <= a) return;
    PROF_START(kKTerminal);
    if (fast && p == 0)
      hipLaunchKernelGGL((terminal_fast_kernel<V, 0>), dim3(b - a), dim3(FT), lds_term, st, d, a, span, B, (const V*)reach,
                         (const uint32_t*)alive, (V*)vals, rowlen);
    else if (fast)
      hipLaunchKernelGGL((terminal_fast_kernel<V, 1>), dim3(b - a), dim3(FT), lds_term, st, d, a, span, B, (const V*)reach,
                         (const uint32_t*)alive, (V*)vals, rowlen);
    else
      hipLaunchKernelGGL((terminal_kernel<V>), dim3(b - a), dim3(TERM_T), lds_term, st, d, p, a, span, B, (const V*)reach,
                         (const uint32_t*)alive, (V*)vals, rowlen);
    PROF_STOP();
  };
  auto up = [&](uint32_t ca, uint32_t cb, uint32_t a, uint32_t b, const Bias& B) {
    if (cb > ca) {
      PROF_START(kKUp);
      hipLaunchKernelGGL((chance_kernel<V>), dim3(cb - ca, tiles), dim3(UP_T), 0, st, d, p, ca, span, B, (V*)vals, rowlen,
                         (const uint32_t*)alive);
      PROF_STOP();
    }
    if (b <= a) return;
    // the walks of one storage pass over the segments [a, b)
    auto walks = [&](auto qp) {
      constexpr int QP = decltype(qp)::value;
      if constexpr (std::is_same<V, float>::value || UMODE == kBest) {
        // f32 walks and the f64 best response, each over a list of the
        // segments among [a, b): those of up to REG_T members (or, when more
        // than a quarter of the walks of up to REG_S members are larger, of
        // up to REG_S) in registers, then those of up to REG_S, REG_M and
        // REG_L, then the larger ones in LDS
        auto cut = [&](const uint32_t* l, uint32_t n, uint32_t& i0, uint32_t& i1) {
          i0 = (uint32_t)(std::lower_bound(l, l + n, a) - l);
          i1 = (uint32_t)(std::lower_bound(l, l + n, b) - l);
        };
        uint32_t i0, i1, j0, j1;
        cut(d.m48_host, d.num_m48, i0, i1);
        cut(d.w48_host, d.num_w48, j0, j1);
        // (f64 walks keep the two kernels: the dual kernel's f64 form spills)
        const bool wide = sizeof(V) == 4 && (uint64_t)(i1 - i0) * 4 > (uint64_t)(j1 - j0);
        if (wide && j1 > j0) {
          PROF_START(kKUp);
          hipLaunchKernelGGL((up_reg_dual_kernel<V, UMODE, QP>), dim3(j1 - j0, tiles), dim3(UP_T), 0, st, d, p, j0, span, B,
                             (V*)vals, rowlen, (const uint32_t*)alive, sto, prm, qout, d.w48);
          PROF_STOP();
        }
        cut(d.w32_host, d.num_w32, j0, j1);
        if (!wide && j1 > j0) {
          PROF_START(kKUp);
          hipLaunchKernelGGL((up_reg_kernel<V, UMODE, (int)REG_T, QP>), dim3(j1 - j0, tiles), dim3(UP_T), 0, st, d, p, j0, span, B,
                             (V*)vals, rowlen, (const uint32_t*)alive, sto, prm, qout, d.w32);
          PROF_STOP();
        }
        if (!wide && i1 > i0) {
          PROF_START(kKUp);
          hipLaunchKernelGGL((up_reg_kernel<V, UMODE, (int)REG_S, QP>), dim3(i1 - i0, tiles), dim3(UP_T), 0, st, d, p, i0, span, B,
                             (V*)vals, rowlen, (const uint32_t*)alive, sto, prm, qout, d.m48);
          PROF_STOP();
        }
        cut(d.mid_host, d.num_mid, i0, i1);
        if (i1 > i0) {
          PROF_START(kKUp);
          hipLaunchKernelGGL((up_reg_kernel<V, UMODE, (int)REG_M, QP>), dim3(i1 - i0, tiles), dim3(UP_T), 0, st, d, p, i0, span, B,
                             (V*)vals, rowlen, (const uint32_t*)alive, sto, prm, qout, d.mid);
          PROF_STOP();
        }
        cut(d.m80_host, d.num_m80, i0, i1);
        if (i1 > i0) {
          PROF_START(kKUp);
          hipLaunchKernelGGL((up_reg_kernel<V, UMODE, (int)REG_L, QP>), dim3(i1 - i0, tiles), dim3(UP_T), 0, st, d, p, i0, span, B,
                             (V*)vals, rowlen, (const uint32_t*)alive, sto, prm, qout, d.m80);
          PROF_STOP();
        }
        cut(d.big_host, d.num_big, i0, i1);
        if (i1 > i0) {
          PROF_START(kKUp);
          hipLaunchKernelGGL((up_kernel<V, UMODE, QP>), dim3(i1 - i0, tiles), dim3(UP_T), lds_up, st, d, p, i0, span, B, (V*)vals,
                             rowlen, (const uint32_t*)alive, sto, prm, qout, d.big);
          PROF_STOP();
        }
        // the shape-specialised walks (the f32 DCFR walk or, with 16-bit
        // storage, its two passes, and the FP64 best response; d's generic
        // lists then leave their segments out)
        constexpr int kind = UMODE == kBest ? jit::kWalkBest
                             : QP == 1    ? jit::kWalkMax16
                             : QP == 2    ? jit::kWalkEnc16
                                          : jit::kWalkF32;
        if constexpr ((std::is_same<V, float>::value && UMODE == kCfr && QP != 3) || (UMODE == kBest && (QP == 0 || QP == 3))) {
          if (jit != nullptr)
            for (size_t k = 0; k < jit->lists.size(); k++) {
              const std::vector<uint32_t>& L = jit->lists[k];
              cut(L.data(), (uint32_t)L.size(), i0, i1);
              if (i1 <= i0) continue;
              hipFunction_t fn = jit->fn[jit::walk_fn(k, kind, p)];
              if (fn == nullptr) continue;  // (not compiled for this storage: the engine never asks for it)
              jit::WalkArgs A{jit->dev[k], d.seg_head, (const unsigned char*)d.mrec, d.malive, (float*)vals, sto.r,
                              (const float*)prm, jit->zrow, sto.rq, sto.rsc, sto.rnx, i0, i1 - i0, tiles, rowlen, B.v0, B.vb};
              void* args[] = {&A};
              PROF_START(kKUp);
              (void)hipModuleLaunchKernel(fn, jit::walk_groups(i1 - i0, tiles), 1, 1, UP_T, 1, 1, 0, st, args, nullptr);
              PROF_STOP();
            }
        }
      } else {
        PROF_START(kKUp);
        hipLaunchKernelGGL((up_kernel<V, UMODE, QP>), dim3(b - a, tiles), dim3(UP_T), lds_up, st, d, p, a, span, B, (V*)vals, rowlen,
                           (const uint32_t*)alive, sto, prm, qout, (const uint32_t*)nullptr);
        PROF_STOP();
      }
    };
    if constexpr (!Q) {
      walks(std::integral_constant<int, 0>());
    } else if constexpr (UMODE == kCfr) {
      walks(std::integral_constant<int, 1>());
      walks(std::integral_constant<int, 2>());
    } else {
      walks(std::integral_constant<int, 3>());
    }
  };
  if (first_depth == 0) down(P.depth_off[0], P.depth_off[1], B0);
  if (leaf_hook && first_depth == 0) (*leaf_hook)();  // leaves read the top segment's reach
  // fast terminals launch over runs (TermBatch), exact ones over records
  auto tlo = [&](uint32_t c) { return fast ? P.tb_coff[c] : P.term_coff[c]; };
  if (C == 0) {  // every row resident (a river tree, or a batch of spots): all depths at once
    for (uint32_t dd = std::max(1u, first_depth); dd < D; dd++) down(P.depth_off[dd], P.depth_off[dd + 1], B0);
    terminals(0, fast ? (uint32_t)P.tbatch.size() : (uint32_t)P.terminals.size(), B0);
    for (uint32_t dd = D; dd-- > std::max(1u, first_depth);)
      up(P.chance_off[dd], P.chance_off[dd + 1], P.depth_off[dd], P.depth_off[dd + 1], B0);
  } else {
    const uint32_t qc = first_depth == 0 ? kNone : P.chunk_of[span.root];
    bool first = true;
    for (const ChunkGroup& g : groups) {
      if (qc != kNone && (qc < g.c0 || qc >= g.c1)) continue;
      const Bias B{P.slot_s0[p], P.slot_coff[p][g.c0] - P.slot_s0[p], P.val_v0, P.val_coff[g.c0] - P.val_v0};
      const uint32_t* so = P.seg_coff.data();
      const uint32_t* co = P.ch_coff.data();
      for (uint32_t dd = std::max(1u, first_depth); dd < D; dd++) down(so[dd * (C + 1) + g.c0], so[dd * (C + 1) + g.c1], B);
      // the top segment's terminals (resident rows) ride with the first group
      terminals(first && first_depth == 0 ? 0u : tlo(g.c0), tlo(g.c1), B);
      first = false;
      for (uint32_t dd = D; dd-- > std::max(1u, first_depth);)
        up(co[dd * (C + 1) + g.c0], co[dd * (C + 1) + g.c1], so[dd * (C + 1) + g.c0], so[dd * (C + 1) + g.c1], B);
    }
  }
  if (first_depth == 0) up(P.chance_off[0], P.chance_off[1], P.depth_off[0], P.depth_off[1], B0);
}

Status Engine::launch_traversal(int p, Mode mode, bool f64, ProfRec* prof) {
  const Plan& P = plan_;
  const DPlan& d = dp_->d;
  const int o = p ^ 1;
  // whole tree: the root segments' down pass starts every spot's reach row
  // and their up pass keeps every spot's root value row of p (after the
  // iteration parameters, where root_rows reads them)
  Span span{0, 1, (uint32_t)P.num_nodes, 0, 1, (char*)params_ + sizeof(IterParams)};
  (void)o;
  const int x = f64 ? 1 : 0;
  void* reach = reach_[x];
  void* vals = vals_[x];
  uint32_t* alive = alive_[x];
  std::function<void()> hook;
  if (!P.leaves.empty()) hook = [&, p, mode, f64]() { launch_leaves(p, mode, f64, reach, vals, alive); };
  const std::function<void()>* hk = P.leaves.empty() ? nullptr : &hook;
  const Store sto = store();
  auto run = [&](auto qc) {
    constexpr bool Q = decltype(qc)::value;
    if (mode == kCfr)
      launch_span<float, kCfr, kCfr, Q>(dp_->dcfr, P, stream_, p, span, 0, groups_[0], reach, alive, sto, params_, vals, rowlen_,
                                        lds_down_, lds_term_[0], lds_up_[0], (float*)nullptr, precision_ == kFast, down_threads_,
                                        prof, hk, &dp_->jit);
    else if (mode == kBest)
      launch_span<double, kAverage, kBest, Q>(dp_->dcfr, P, stream_, p, span, 0, groups_[1], reach, alive, sto, params_, vals,
                                              rowlen_, lds_down_, lds_term_[1], lds_up_[1], (double*)nullptr, precision_ == kFast,
                                              down_threads_, nullptr, hk, &dp_->jit);
    else if (f64)
      launch_span<double, kAverage, kAverage, Q>(d, P, stream_, p, span, 0, groups_[1], reach, alive, sto, params_, vals, rowlen_,
                                                 lds_down_, lds_term_[1], lds_up_[1], (double*)nullptr, precision_ == kFast,
                                                 down_threads_, nullptr, hk);
    else
      launch_span<float, kAverage, kAverage, Q>(d, P, stream_, p, span, 0, groups_[0], reach, alive, sto, params_, vals, rowlen_,
                                                lds_down_, lds_term_[0], lds_up_[0], (float*)nullptr, precision_ == kFast,
                                                down_threads_, nullptr, hk);
  };
  if (q16_) run(std::true_type());
  else run(std::false_type());
  if (q16_ && mode == kCfr && P.num_pslots > 0)
    hipLaunchKernelGGL(commit_scales_kernel, dim3((P.num_pslots + 255) / 256), dim3(256), 0, stream_, sto, P.num_pslots);
  HIPCHK(hipGetLastError());
  return kOk;
}

// The leaves of a depth-limited tree for traverser p (after the top
// segment's down pass): zeros without a network.
void Engine::launch_leaves(int p, Mode mode, bool f64, void* reach, void* vals, uint32_t* alive) {
  const Plan& P = plan_;
  const DPlan& d = dp_->d;
  const uint32_t nl = (uint32_t)P.leaves.size();
  const Bias B0{P.slot_s0[p], 0, P.val_v0, 0};
  if (!net_) {
    if (f64) hipLaunchKernelGGL((leaf_zero_kernel<double>), dim3(nl), dim3(256), 0, stream_, B0, leaves_, (double*)vals, rowlen_);
    else hipLaunchKernelGGL((leaf_zero_kernel<float>), dim3(nl), dim3(256), 0, stream_, B0, leaves_, (float*)vals, rowlen_);
    return;
  }
  const uint32_t seg = P.seg_of[P.leaves[0].node];
  if (f64) {
    if (mode == kCfr)
      hipLaunchKernelGGL((leaf_own_kernel<double, false>), dim3(1), dim3(1024), 0, stream_, d, p, seg, (const float*)regrets_,
                         (const float*)sums_, (double*)mreach_, leaves_, nl, (double*)own_, rowlen_);
    else
      hipLaunchKernelGGL((leaf_own_kernel<double, true>), dim3(1), dim3(1024), 0, stream_, d, p, seg, (const float*)regrets_,
                         (const float*)sums_, (double*)mreach_, leaves_, nl, (double*)own_, rowlen_);
    hipLaunchKernelGGL((leaf_input_kernel<double>), dim3(net_rows_), dim3(256), 0, stream_, d, p, B0, leaves_, nl, deal_, ncards_,
                       board_, board_len_, (const double*)own_, (const double*)reach, rowlen_, net_in_);
  } else {
    if (mode == kCfr)
      hipLaunchKernelGGL((leaf_own_kernel<float, false>), dim3(1), dim3(1024), 0, stream_, d, p, seg, (const float*)regrets_,
                         (const float*)sums_, (float*)mreach_, leaves_, nl, (float*)own_, rowlen_);
    else
      hipLaunchKernelGGL((leaf_own_kernel<float, true>), dim3(1), dim3(1024), 0, stream_, d, p, seg, (const float*)regrets_,
                         (const float*)sums_, (float*)mreach_, leaves_, nl, (float*)own_, rowlen_);
    hipLaunchKernelGGL((leaf_input_kernel<float>), dim3(net_rows_), dim3(256), 0, stream_, d, p, B0, leaves_, nl, deal_, ncards_,
                       board_, board_len_, (const float*)own_, (const float*)reach, rowlen_, net_in_);
  }
  (void)net_->forward_device(net_in_, net_rows_, net_out_, stream_);
  if (f64)
    hipLaunchKernelGGL((leaf_value_kernel<double>), dim3(nl), dim3(256), 0, stream_, d, p, B0, leaves_, deal_, ncards_,
                       (const double*)reach, (const uint32_t*)alive, (const float*)net_out_, (double*)vals, rowlen_);
  else
    hipLaunchKernelGGL((leaf_value_kernel<float>), dim3(nl), dim3(256), 0, stream_, d, p, B0, leaves_, deal_, ncards_,
                       (const float*)reach, (const uint32_t*)alive, (const float*)net_out_, (float*)vals, rowlen_);
}

Status Engine::attach_net(net::Net* n, const std::vector<int>& board) {
  const Plan& P = plan_;
  if (P.leaves.empty()) {
    set_error("invalid input: the tree has no leaves for a value network");
    return kInvalidInput;
  }
  if (q16_) {
    set_error("invalid input: value-network leaves need f32 storage");
    return kInvalidInput;
  }
  if (P.num_spots != 1) {
    set_error("invalid input: value-network leaves run on single-spot engines");
    return kInvalidInput;
  }
  std::vector<int8_t> deal, bd;
  uint64_t used = 0;
  for (int c : board) {
    used |= 1ull << c;
    bd.push_back((int8_t)c);
  }
  for (int c = 0; c < 52; c++)
    if (!(used >> c & 1)) deal.push_back((int8_t)c);
  ncards_ = (uint32_t)deal.size();
  board_len_ = (uint32_t)bd.size();
  const uint32_t rows = (uint32_t)P.leaves.size() * ncards_;
  net_rows_ = (rows + 63) / 64 * 64;
  if (n && net_rows_ > n->capacity()) {
    set_error("invalid input: the network's batch capacity is below the leaves' rows");
    return kInvalidInput;
  }
  if (!leaves_) {
    DeviceBuf b;
    Status st;
    const size_t nl = P.leaves.size();
    if ((st = alloc(b, nl * sizeof(LeafRec) + 64 + 64 + (size_t)nl * rowlen_ * 8 + (size_t)P.max_members * rowlen_ * 8 +
                           (size_t)net_rows_ * 2752 * 2 + (size_t)net_rows_ * 2652 * 4 + 1024)) != kOk)
      return st;
    char* q = (char*)b.p;
    auto take = [&](size_t bytes) {
      char* r = q;
      q += (bytes + 255) & ~size_t(255);
      return r;
    };
    leaves_ = (const LeafRec*)take(nl * sizeof(LeafRec));
    deal_ = (int8_t*)take(64);
    board_ = (int8_t*)take(64);
    own_ = take((size_t)nl * rowlen_ * 8);
    mreach_ = take((size_t)P.max_members * rowlen_ * 8);
    net_in_ = (__bf16*)take((size_t)net_rows_ * 2752 * 2);
    net_out_ = (float*)take((size_t)net_rows_ * 2652 * 4);
    HIPCHK(hipMemcpy((void*)leaves_, P.leaves.data(), nl * sizeof(LeafRec), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(deal_, deal.data(), deal.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(board_, bd.data(), bd.size(), hipMemcpyHostToDevice));
  net_ = n;
  if (use_graph_) {
    for (int i = 0; i < 3; i++)
      if (graph_[i]) {
        (void)hipGraphExecDestroy(graph_[i]);
        graph_[i] = nullptr;
      }
    return build_graphs();
  }
  return kOk;
}

Status Engine::write_storage(uint64_t offset, size_t len, const float* regrets, const float* sums) {
  if (q16_) {
    set_error("invalid input: 16-bit storage");
    return kInvalidInput;
  }
  if (offset + len > plan_.storage_len) {
    set_error("invalid input: storage range");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  if (regrets) HIPCHK(hipMemcpy(regrets_ + offset, regrets, len * sizeof(float), hipMemcpyHostToDevice));
  if (sums) HIPCHK(hipMemcpy(sums_ + offset, sums, len * sizeof(float), hipMemcpyHostToDevice));
  return kOk;
}

Status Engine::build_graphs() {
  for (int g = 0; g < 3; g++) {
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    Status st = kOk;
    for (int p = 0; p < 2 && st == kOk; p++)
      st = launch_traversal(p, g == 0 ? kCfr : g == 1 ? kBest : kAverage, g != 0);
    hipError_t ec = hipStreamEndCapture(stream_, &graph);
    if (st != kOk) return st;
    HIPCHK(ec);
    HIPCHK(hipGraphInstantiate(&graph_[g], graph, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(graph));
  }
  return kOk;
}

Status Engine::run_graph(int which) {
  if (use_graph_) {
    HIPCHK(hipGraphLaunch(graph_[which], stream_));
    return kOk;
  }
  for (int p = 0; p < 2; p++) {
    Status st = launch_traversal(p, which == 0 ? kCfr : which == 1 ? kBest : kAverage, which != 0);
    if (st != kOk) return st;
  }
  return kOk;
}

Status Engine::cfr_iteration(const IterParams& dsc) {
  *params_host_ = dsc;
  HIPCHK(hipMemcpyAsync(params_, params_host_, sizeof(IterParams), hipMemcpyHostToDevice, stream_));
  Status st = run_graph(0);
  if (st != kOk) return st;
  HIPCHK(hipStreamSynchronize(stream_));
  return kOk;
}

Status Engine::profile_iteration(const IterParams& dsc, Profile& out) {
  *params_host_ = dsc;
  HIPCHK(hipMemcpyAsync(params_, params_host_, sizeof(IterParams), hipMemcpyHostToDevice, stream_));
  ProfRec rec;
  rec.st = stream_;
  for (int p = 0; p < 2; p++) {
    Status st = launch_traversal(p, kCfr, false, &rec);
    if (st != kOk) return st;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  out = Profile();
  for (auto& e : rec.ev) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
    out.ms[e.first] += ms;
    out.launches[e.first] += 1;
  }
  if (!rec.ev.empty()) {
    float w = 0.0f;
    HIPCHK(hipEventElapsedTime(&w, rec.ev.front().second.first, rec.ev.back().second.second));
    out.wall_ms = w;
  }
  return kOk;
}

Status Engine::traverse(int p, Mode mode, const IterParams& dsc, std::vector<double>& out) {
  *params_host_ = dsc;
  HIPCHK(hipMemcpyAsync(params_, params_host_, sizeof(IterParams), hipMemcpyHostToDevice, stream_));
  bool f64 = mode != kCfr;
  Status st = launch_traversal(p, mode, f64);
  if (st != kOk) return st;
  size_t n = plan_.spot(0).n[p];
  out.assign(n, 0.0);
  if (f64) {
    HIPCHK(hipMemcpyAsync(out.data(), (char*)params_ + result_offset(0, p, rowlen_, 8), n * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
  } else {
    std::vector<float> f(n);
    HIPCHK(hipMemcpyAsync(f.data(), (char*)params_ + result_offset(0, p, rowlen_, 4), n * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n; i++) out[i] = f[i];
  }
  return kOk;
}

Status Engine::root_values(Mode mode, double res[2]) {  // cfr.rs:482-490, in FP64
  std::vector<double> all;
  Status st = root_values_all(mode, all);
  if (st != kOk) return st;
  res[0] = all[0];
  res[1] = all[1];
  return kOk;
}

Status Engine::root_rows(Mode mode, std::vector<double>& rows) {
  Status st = run_graph(mode == kBest ? 1 : 2);
  if (st != kOk) return st;
  const uint32_t K = plan_.num_spots;
  rows.assign((size_t)K * 2 * rowlen_, 0.0);
  HIPCHK(hipMemcpyAsync(rows.data(), (char*)params_ + result_offset(0, 0, rowlen_, 8), rows.size() * 8, hipMemcpyDeviceToHost,
                        stream_));
  HIPCHK(hipStreamSynchronize(stream_));
  return kOk;
}

Status Engine::root_values_all(Mode mode, std::vector<double>& res) {
  std::vector<double> rows;
  Status st = root_rows(mode, rows);
  if (st != kOk) return st;
  const uint32_t K = plan_.num_spots;
  res.assign((size_t)K * 2, 0.0);
  for (uint32_t k = 0; k < K; k++) {
    const Plan& T = plan_.spot(k);
    for (int p = 0; p < 2; p++) {
      double s = 0.0;
      const double* row = rows.data() + ((size_t)k * 2 + p) * rowlen_;
      for (uint32_t i = 0; i < T.n[p]; i++) s += row[i] * (double)T.weights[p][i];
      res[(size_t)k * 2 + p] = s;
    }
  }
  return kOk;
}

Status Engine::retune(uint32_t run_length, uint32_t down_threads) {
  if (down_threads != 128 && down_threads != 256 && down_threads != 512 && down_threads != 1024) {
    set_error("invalid input: down-kernel workgroup size");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  build_runs(plan_, run_length);
  std::string err;
  for (int f64 = 0; f64 < 2 && err.empty(); f64++) {
    ScratchRows r = partition_chunks(plan_, f64 ? 8 : 4, rowlen_, budget_, groups_[f64]);
    err = check_groups(plan_, groups_[f64], r);
  }
  if (!err.empty()) {
    set_error("invalid input: " + err);
    return kInvalidInput;
  }
  HIPCHK(hipMemcpy(tbatch_dev_, plan_.tbatch.data(), plan_.tbatch.size() * sizeof(TermBatch), hipMemcpyHostToDevice));
  down_threads_ = down_threads;
  run_length_ = std::min(std::max(run_length, 1u), kBatchMax);
  if (use_graph_) {
    for (int i = 0; i < 3; i++)
      if (graph_[i]) {
        (void)hipGraphExecDestroy(graph_[i]);
        graph_[i] = nullptr;
      }
    Status st = build_graphs();
    if (st != kOk) return st;
  }
  return kOk;
}

Status Engine::read_storage(uint64_t offset, size_t len, float* regrets, float* sums) const {
  if (q16_) {
    set_error("invalid input: 16-bit storage");
    return kInvalidInput;
  }
  if (offset + len > plan_.storage_len) {
    set_error("invalid input: storage range");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  if (regrets) HIPCHK(hipMemcpy(regrets, regrets_ + offset, len * sizeof(float), hipMemcpyDeviceToHost));
  if (sums) HIPCHK(hipMemcpy(sums, sums_ + offset, len * sizeof(float), hipMemcpyDeviceToHost));
  return kOk;
}

Status Engine::zero_storage() {
  HIPCHK(hipStreamSynchronize(stream_));
  return upload(nullptr, nullptr);
}

Status Engine::set_active(const std::vector<uint32_t>& active) {
  if (active.size() != plan_.num_spots) {
    set_error("invalid input: one flag per spot");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  HIPCHK(hipMemcpy(active_, active.data(), active.size() * 4, hipMemcpyHostToDevice));
  return kOk;
}

Status Engine::set_frozen(const std::vector<uint32_t>& nodes, bool on) {
  HIPCHK(hipStreamSynchronize(stream_));
  for (uint32_t id : nodes) {
    if (id >= plan_.num_nodes || plan_.kind[id] > kP1) {
      set_error("invalid input: only decision nodes are frozen");
      return kInvalidInput;
    }
    const size_t m = (size_t)plan_.seg_mem_off[plan_.seg_of[id]] + plan_.local[id];
    if (m >= plan_.mrec.size() || plan_.members[m] != id) {
      set_error("invalid input: node has no member record");
      return kInvalidInput;
    }
    MemRec& r = plan_.mrec[m];
    r.tflags = on ? (uint8_t)(r.tflags | kNodeFrozen) : (uint8_t)(r.tflags & ~kNodeFrozen);
    HIPCHK(hipMemcpy((char*)dp_->d.mrec + m * sizeof(MemRec) + offsetof(MemRec, tflags), &r.tflags, 1,
                     hipMemcpyHostToDevice));
  }
  return kOk;
}

Status Engine::exploitability(bool raked, double* out) {  // cfr.rs:494-502
  double br[2];
  Status st = root_values(kBest, br);
  if (st != kOk) return st;
  if (raked) {
    double ev[2];
    if ((st = root_values(kAverage, ev)) != kOk) return st;
    *out = ((br[0] - ev[0]) + (br[1] - ev[1])) / 2.0;
  } else {
    *out = (br[0] + br[1]) / 2.0;
  }
  return kOk;
}

Status Engine::subtree_values(uint32_t node, int q, const std::vector<float>& opp, std::vector<float>& child_vals) {
  const Plan& P = plan_;
  if (P.num_spots != 1) {
    set_error("invalid input: queries run on single-spot engines");
    return kInvalidInput;
  }
  if (node >= P.num_nodes || (int)P.kind[node] != q) {
    set_error("invalid input: query node is not a decision node of that player");
    return kInvalidInput;
  }
  const int o = q ^ 1;
  if (opp.size() != P.n[o]) {
    set_error("invalid input: reach length");
    return kInvalidInput;
  }
  const uint32_t nq = P.n[o], np = P.n[q], na = P.na[node];
  // stage the reach row and the output in the scratch value area's tail
  DeviceBuf tmp;
  Status st = alloc(tmp, (size_t)nq * 4 + (size_t)na * np * 4);
  if (st != kOk) return st;
  float* dreach = (float*)tmp.p;
  float* dout = dreach + nq;
  HIPCHK(hipMemcpyAsync(dreach, opp.data(), (size_t)nq * 4, hipMemcpyHostToDevice, stream_));
  bool any = false;
  for (float x : opp) any |= x != 0.0f;
  uint32_t slot = P.rslot[(size_t)q * P.num_nodes + node];
  const uint32_t qc = P.chunk_of[node];
  if (slot >= P.slot_s0[q] && qc != kNone)
    for (const ChunkGroup& g : groups_[0])
      if (qc >= g.c0 && qc < g.c1) slot -= P.slot_coff[q][g.c0] - P.slot_s0[q];
  hipLaunchKernelGGL((begin_row_kernel<float>), dim3((nq + 255) / 256), dim3(256), 0, stream_, nq, (const float*)dreach,
                     (float*)reach_[0], slot, alive_[0], any ? 1u : 0u, rowlen_);
  Span span{node, P.first_child[node], P.subtree_end[node], P.seg_of[node]};
  span.kid0 = P.first_child[node];
  span.kid1 = P.first_child[node] + na;
  // depths from the node's segment on
  uint32_t seg = P.seg_of[node], first_depth = 0;
  while (first_depth + 1 < P.depth_off.size() && P.depth_off[first_depth + 1] <= seg) first_depth++;
  if (q16_)
    launch_span<float, kAverage, kAverage, true>(dp_->d, P, stream_, q, span, first_depth, groups_[0], reach_[0], alive_[0], store(),
                                                 params_, vals_[0], rowlen_, lds_down_, lds_term_[0], lds_up_[0], dout,
                                                 precision_ == kFast, down_threads_);
  else
    launch_span<float, kAverage, kAverage, false>(dp_->d, P, stream_, q, span, first_depth, groups_[0], reach_[0], alive_[0], store(),
                                                  params_, vals_[0], rowlen_, lds_down_, lds_term_[0], lds_up_[0], dout,
                                                  precision_ == kFast, down_threads_);
  HIPCHK(hipGetLastError());
  child_vals.assign((size_t)na * np, 0.0f);
  HIPCHK(hipMemcpyAsync(child_vals.data(), dout, (size_t)na * np * 4, hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipStreamSynchronize(stream_));
  (void)hipFree(tmp.p);
  bufs_.pop_back();
  bytes_ -= tmp.bytes;
  return kOk;
}

}  // namespace pts
