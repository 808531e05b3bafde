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

#include "engine.hpp"

namespace pts {

static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const std::string& last_error() { return g_err; }

#define HIPCHK(x)                                                                                  \
  do {                                                                                             \
    hipError_t e_ = (x);                                                                           \
    if (e_ != hipSuccess) {                                                                        \
      set_error(std::string("HIP error in ") + #x + ": " + hipGetErrorString(e_));                 \
      return e_ == hipErrorOutOfMemory ? kOutOfMemory : kDriver;                                   \
    }                                                                                              \
  } while (0)

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint16_t NONE16 = 0xFFFF;
constexpr int MAXA = 8;       // actions per decision node (checked at create)
constexpr int TERM_T = 64;    // terminal workgroup: one wave
constexpr int UP_T = 64;      // up workgroup: one wave, one hand per lane
constexpr int DOWN_T = 256;   // down workgroup
constexpr int MAX_K = 21;     // ceil(1326 / 64): hands per lane in the terminal kernel

// Device view of the plan (all pointers into one HBM blob).
struct DPlan {
  uint32_t num_nodes;
  uint32_t n[2];
  uint32_t max_entries;
  double half_pot, norm;
  const uint8_t *kind, *na, *tflags;
  const uint32_t *first_child, *vslot, *tboard, *rslot;
  const uint64_t* offset;
  const int32_t* invest;
  const float *rake, *factor;
  const int8_t* card;
  const uint16_t *local, *alive_src;
  const uint32_t *seg_root, *seg_parent, *seg_mem_off, *seg_mem_cnt, *members, *terminals;
  const uint8_t* cards[2];
  const uint32_t* same[2];
  const uint32_t* bl_off[2];
  const uint16_t* bl_idx[2];
  const uint32_t *ent_off[2], *ent_cnt[2], *cl_base[2];
  const uint16_t *ent_idx[2], *cl_off[2], *cl_idx[2], *my_cnt[2], *hand_ent[2];
};

// The nodes a traversal covers: {root} and [lo, hi) (the whole tree, or the
// subtree below a queried decision node).
struct Span {
  uint32_t root, lo, hi, seg;  // seg: segment holding `root`
  __device__ bool has(uint32_t id) const { return id == root || (id >= lo && id < hi); }
};

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

// ------------------------------------------------------------------ begin
template <typename V>
__global__ void begin_kernel(uint32_t nq, const float* __restrict__ w, V* __restrict__ reach, uint32_t slot,
                             uint32_t* __restrict__ alive, uint32_t is_alive) {
  uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h < nq) reach[(size_t)slot * nq + h] = (V)w[h];
  if (h == 0) alive[slot] = is_alive;
}

// ------------------------------------------------------------------ down
// One workgroup per segment: the opponent's reach pushed down the segment's
// members in breadth-first order. CFR mode (V = float): regret matching and
// the strategy-sum update (cfr.rs:156-207,389-419; river.rs:116-160).
// Otherwise the normalized average strategy (cfr.rs:124-154).
template <typename V, int MODE>
__global__ void __launch_bounds__(DOWN_T) down_kernel(DPlan P, int p, uint32_t seg0, Span span, V* __restrict__ reach,
                                                      uint32_t* __restrict__ alive, const float* __restrict__ regrets,
                                                      float* __restrict__ sums, const IterParams* __restrict__ prm) {
  extern __shared__ uint8_t m_alive[];
  __shared__ uint32_t red;
  const uint32_t seg = seg0 + blockIdx.x;
  const int o = p ^ 1;
  const uint32_t nq = P.n[o], N = P.num_nodes, tid = threadIdx.x;
  const uint32_t* rs = P.rslot + (size_t)p * N;
  const uint16_t* src = P.alive_src + (size_t)p * N;
  const uint32_t root = P.seg_root[seg];
  if (!span.has(root) && seg != span.seg) return;
  if (seg == span.seg) {
    if (tid == 0) m_alive[src[span.root]] = (uint8_t)(alive[rs[span.root]] != 0);
  } else {
    // Chance child: the parent's reach times the chance factor, opponent
    // hands holding the dealt card zeroed (cfr.rs:449-456).
    const uint32_t par = P.seg_parent[seg], sp = rs[par], sm = rs[root];
    uint32_t any = 0;
    if (alive[sp]) {
      const V f = (V)P.factor[par];
      const int c = P.card[root];
      const V* pr = reach + (size_t)sp * nq;
      V* mr = reach + (size_t)sm * nq;
      for (uint32_t h = tid; h < nq; h += DOWN_T) {
        V v = pr[h] * f;
        if (holds(P.cards[o], h, c)) v = V(0);
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
  const uint32_t moff = P.seg_mem_off[seg], mcnt = P.seg_mem_cnt[seg];
  float g = 0.0f;
  bool weighted = false;
  if (MODE == kCfr) {
    g = prm->strat;
    weighted = prm->reach_weighted != 0;
  }
  for (uint32_t m = 0; m < mcnt; m++) {
    const uint32_t node = P.members[moff + m];
    if ((int)P.kind[node] != o || !span.has(node)) continue;
    const uint32_t na = P.na[node], fc = P.first_child[node];
    if (!m_alive[src[node]]) {
      // A dead node's children are dead (cfr.rs:271-274, river.rs:103-107).
      if (tid < na) {
        alive[rs[fc + tid]] = 0;
        m_alive[P.local[fc + tid]] = 0;
      }
      __syncthreads();
      continue;
    }
    const V* rrow = reach + (size_t)rs[node] * nq;
    const uint64_t off = P.offset[node];
    uint32_t bits = 0;
    for (uint32_t h = tid; h < nq; h += DOWN_T) {
      const V r = rrow[h];
      V sa[MAXA];
      if (MODE == kCfr) {
        float rg[MAXA];
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) rg[a] = regrets[off + (size_t)a * nq + h];
        float d = pos0(rg[0]);
#pragma unroll
        for (int a = 1; a < MAXA; a++)
          if (a < (int)na) d += pos0(rg[a]);
        const float inv = recip(d), uni = 1.0f / (float)na;
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) {
            const float s = inv == 0.0f ? uni : pos0(rg[a]) * inv;
            float* x = sums + off + (size_t)a * nq + h;
            *x = weighted ? *x * g + s * (float)r : *x * g + s;
            sa[a] = (V)s;
          }
      } else {
        V sv[MAXA];
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) sv[a] = (V)sums[off + (size_t)a * nq + h];
        V d = sv[0];
#pragma unroll
        for (int a = 1; a < MAXA; a++)
          if (a < (int)na) d += sv[a];
        const V inv = recip(d), uni = V(1) / (V)na;
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) sa[a] = inv == V(0) ? uni : sv[a] * inv;
      }
#pragma unroll
      for (int a = 0; a < MAXA; a++)
        if (a < (int)na) {
          const V v = sa[a] * r;
          reach[(size_t)rs[fc + a] * nq + h] = v;
          if (v != V(0)) bits |= 1u << a;
        }
    }
    bits = block_or(bits, &red);
    if (tid < na) {
      const uint32_t b = (bits >> tid) & 1u;
      alive[rs[fc + tid]] = b;
      m_alive[P.local[fc + tid]] = (uint8_t)b;
    }
    __syncthreads();
  }
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
template <typename V>
__global__ void __launch_bounds__(TERM_T) terminal_kernel(DPlan P, int p, Span span, const V* __restrict__ reach,
                                                          const uint32_t* __restrict__ alive, V* __restrict__ vals,
                                                          uint32_t rowlen) {
  extern __shared__ double sh[];
  const uint32_t t = P.terminals[blockIdx.x];
  if (!span.has(t)) return;
  const uint32_t N = P.num_nodes;
  const uint32_t s = P.rslot[(size_t)p * N + t];
  if (!alive[s]) return;
  const int o = p ^ 1;
  const uint32_t nq = P.n[o], np = P.n[p], tid = threadIdx.x;
  const uint32_t mmax = P.max_entries;
  double* Pa = sh;                                  // [mmax]
  double* Ca = sh + mmax;                           // [max(2*mmax, 53)]
  V* rv = (V*)(Ca + (2 * mmax > 53 ? 2 * mmax : 53));  // [nq]
  const V* R = reach + (size_t)s * nq;
  V* out = vals + (size_t)P.vslot[t] * rowlen;
  for (uint32_t h = tid; h < nq; h += TERM_T) rv[h] = R[h];
  __syncthreads();
  const uint8_t k = P.kind[t], fl = P.tflags[t];
  const bool batched = (fl & kTermBatched) != 0;
  const double hp = P.half_pot, norm = P.norm, rake = (double)P.rake[t];
  const int32_t inv_p = P.invest[2 * t + p], inv_o = P.invest[2 * t + o];
  const uint8_t* cp = P.cards[p];
  const uint8_t* co = P.cards[o];
  const uint32_t* same_p = P.same[p];

  if (k != kShowdown) {
    const int f = k - kFold0;
    // cfr.rs:224-235
    const double pay = (f == p ? -(hp + (double)inv_p) : hp + (double)P.invest[2 * t + f] - rake) * norm;
    const bool paired = (fl & (p == 0 ? kTermPaired0 : kTermPaired1)) != 0;
    if (tid <= 52) {
      double acc = 0.0;
      if (paired) {
        const uint32_t b = P.tboard[t];
        if (tid == 52) {
          const uint32_t eo = P.ent_off[o][b], cnt = P.ent_cnt[o][b];
          for (uint32_t j = 0; j < cnt; j++) acc += (double)rv[P.ent_idx[o][eo + j]];
        } else {
          const uint16_t* clo = P.cl_off[o] + (size_t)b * 53;
          const uint16_t* ci = P.cl_idx[o] + P.cl_base[o][b];
          for (uint32_t x = clo[tid]; x < clo[tid + 1]; x++) acc += (double)rv[ci[x]];
        }
      } else {
        if (tid == 52) {
          for (uint32_t h = 0; h < nq; h++) acc += (double)rv[h];
        } else {
          const uint16_t* bi = P.bl_idx[o];
          for (uint32_t x = P.bl_off[o][tid]; x < P.bl_off[o][tid + 1]; x++) acc += (double)rv[bi[x]];
        }
      }
      Ca[tid] = acc;
    }
    __syncthreads();
    const double total = Ca[52];
    for (uint32_t i = tid; i < np; i += TERM_T) {
      const int a = cp[2 * i], b = cp[2 * i + 1];
      const uint32_t sm = same_p[i];
      const double sw = sm == NONE ? 0.0 : (double)rv[sm];
      V v;
      if (batched) v = (V)(pay * (total - Ca[a] - Ca[b] + sw));
      else v = total == 0.0 ? V(0) : (V)(pay * (total - (Ca[a] + Ca[b]) + sw));
      out[i] = v;
    }
    return;
  }

  // showdown (cfr.rs:238-246 payoffs)
  const double win = (hp + (double)inv_o - rake) * norm;
  const double lose = -(hp + (double)inv_p) * norm;
  const double tie = 0.5 * ((double)(inv_o - inv_p) - rake) * norm;
  const uint32_t b = P.tboard[t];
  const uint32_t eo = P.ent_off[o][b], mo = P.ent_cnt[o][b];
  const uint16_t* eidx = P.ent_idx[o] + eo;
  const uint16_t* clo = P.cl_off[o] + (size_t)b * 53;
  const uint16_t* ci = P.cl_idx[o] + P.cl_base[o][b];
  const uint32_t ep = P.ent_off[p][b];
  const uint16_t* hent = P.hand_ent[p] + (size_t)b * np;
  const uint16_t* c6 = P.my_cnt[p];

  // ascending sweep: weaker opponent hands
  if (tid == 52) {
    double acc = 0.0;
    for (uint32_t j = 0; j < mo; j++) { acc += (double)rv[eidx[j]]; Pa[j] = acc; }
  } else if (tid < 52) {
    double acc = 0.0;
    for (uint32_t x = clo[tid]; x < clo[tid + 1]; x++) { acc += (double)rv[ci[x]]; Ca[x] = acc; }
  }
  __syncthreads();
  double W[MAX_K];
#pragma unroll
  for (int kk = 0; kk < MAX_K; kk++) {
    W[kk] = 0.0;
    const uint32_t i = tid + kk * TERM_T;
    if (i < np) {
      const uint16_t e = hent[i];
      if (e != NONE16) {
        const uint16_t* c = c6 + (size_t)(ep + e) * 6;
        const int c1 = cp[2 * i], c2 = cp[2 * i + 1];
        const double Pv = c[0] ? Pa[c[0] - 1] : 0.0;
        const double P1 = c[2] ? Ca[clo[c1] + c[2] - 1] : 0.0;
        const double P2 = c[4] ? Ca[clo[c2] + c[4] - 1] : 0.0;
        W[kk] = batched ? win * (Pv - P1 - P2) : win * (Pv - (P1 + P2));
      }
    }
  }
  __syncthreads();
  // descending sweep: stronger opponent hands
  if (tid == 52) {
    double acc = 0.0;
    for (uint32_t j = mo; j-- > 0;) { acc += (double)rv[eidx[j]]; Pa[j] = acc; }
  } else if (tid < 52) {
    double acc = 0.0;
    for (uint32_t x = clo[tid + 1]; x-- > clo[tid];) { acc += (double)rv[ci[x]]; Ca[x] = acc; }
  }
  __syncthreads();
#pragma unroll
  for (int kk = 0; kk < MAX_K; kk++) {
    const uint32_t i = tid + kk * TERM_T;
    if (i < np) {
      V v = V(0);
      const uint16_t e = hent[i];
      if (e != NONE16) {
        const uint16_t* c = c6 + (size_t)(ep + e) * 6;
        const int c1 = cp[2 * i], c2 = cp[2 * i + 1];
        const uint32_t len1 = clo[c1 + 1] - clo[c1], len2 = clo[c2 + 1] - clo[c2];
        const double Sv = c[1] < mo ? Pa[c[1]] : 0.0;
        const double S1 = c[3] < len1 ? Ca[clo[c1] + c[3]] : 0.0;
        const double S2 = c[5] < len2 ? Ca[clo[c2] + c[5]] : 0.0;
        if (batched) {
          const double r = W[kk] + lose * (Sv - S1 - S2);
          v = (V)r;
        } else {
          v = (V)W[kk];
          v = v + (V)(lose * (Sv - (S1 + S2)));
        }
        if (tie != 0.0) {  // eval.rs:347-379
          double es = 0.0, m1 = 0.0, m2 = 0.0;
          for (uint32_t j = c[0]; j < c[1]; j++) {
            const uint32_t idx = eidx[j];
            const double w = (double)rv[idx];
            es += w;
            if (holds(co, idx, c1)) m1 += w;
            if (holds(co, idx, c2)) m2 += w;
          }
          const uint32_t sm = same_p[i];
          const double sw = sm == NONE ? 0.0 : (double)rv[sm];
          const double eq = es - (m1 + m2) + sw;
          v = v + (V)(tie * eq);
        }
      }
      out[i] = v;
    }
  }
}

// ------------------------------------------------------------------ up
// One wave per (segment, 64-hand tile), one hand per lane: values pulled up
// the segment's members in reverse breadth-first order (no barriers: every
// lane owns its hand). Own nodes: cfr.rs:315-382 (regret update in CFR
// mode); opponent nodes: sum of the live children (cfr.rs:421-432,
// river.rs:264-271); chance nodes: f64 sum over dealt cards, own hands
// holding the card skipped (cfr.rs:457-478).
template <typename V, int MODE>
__global__ void __launch_bounds__(UP_T) up_kernel(DPlan P, int p, uint32_t seg0, Span span, V* __restrict__ vals,
                                                  uint32_t rowlen, const uint32_t* __restrict__ alive,
                                                  float* __restrict__ regrets, const float* __restrict__ sums,
                                                  const IterParams* __restrict__ prm, V* __restrict__ qout) {
  extern __shared__ unsigned char lds_raw[];
  V* lv = (V*)lds_raw;
  const uint32_t seg = seg0 + blockIdx.x;
  const uint32_t np = P.n[p], N = P.num_nodes, tid = threadIdx.x;
  const uint32_t h = blockIdx.y * UP_T + tid;
  const uint32_t root = P.seg_root[seg];
  if (!span.has(root) && seg != span.seg) return;
  if (h >= np) return;
  const uint32_t* rs = P.rslot + (size_t)p * N;
  const uint32_t moff = P.seg_mem_off[seg], mcnt = P.seg_mem_cnt[seg];
  const uint8_t* cp = P.cards[p];
  float pos = 0.0f, neg = 0.0f;
  if (MODE == kCfr) {
    pos = prm->pos;
    neg = prm->neg;
  }
  for (uint32_t m = mcnt; m-- > 0;) {
    const uint32_t node = P.members[moff + m];
    if (!span.has(node)) continue;
    const uint8_t k = P.kind[node];
    const uint32_t na = P.na[node], fc = P.first_child[node];
    if (qout != nullptr && node == span.root) {
      // query: the values of the decision node's children (lib.rs:854-863)
      for (uint32_t a = 0; a < na; a++) qout[(size_t)a * np + h] = lv[(size_t)P.local[fc + a] * UP_T + tid];
      return;
    }
    const bool live = alive[rs[node]] != 0;
    V v = V(0);
    if (k >= kFold0) {
      if (live) v = vals[(size_t)P.vslot[node] * rowlen + h];
    } else if (k == kChance) {
      if (live) {
        double acc = 0.0;
        for (uint32_t a = 0; a < na; a++) {
          const uint32_t c = fc + a;
          if (holds(cp, h, P.card[c]) || !alive[rs[c]]) continue;
          acc += (double)vals[(size_t)P.vslot[c] * rowlen + h];
        }
        v = (V)acc;
      }
    } else if ((int)k == p) {
      if (live) {
        V cv[MAXA];
#pragma unroll
        for (int a = 0; a < MAXA; a++)
          if (a < (int)na) cv[a] = lv[(size_t)P.local[fc + a] * UP_T + tid];
        const uint64_t off = P.offset[node];
        if (MODE == kCfr) {
          float rg[MAXA];
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) rg[a] = regrets[off + (size_t)a * np + h];
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
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) {
              const float x = rg[a];
              const float disc = x > 0.0f ? pos : neg;
              regrets[off + (size_t)a * np + h] = x * disc + (float)cv[a] - outv;
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
#pragma unroll
          for (int a = 0; a < MAXA; a++)
            if (a < (int)na) sv[a] = (V)sums[off + (size_t)a * np + h];
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
      }
    } else {
      if (live) {
        for (uint32_t a = 0; a < na; a++)
          if (alive[rs[fc + a]]) v += lv[(size_t)P.local[fc + a] * UP_T + tid];
      }
    }
    lv[(size_t)m * UP_T + tid] = v;
  }
  vals[(size_t)P.vslot[root] * rowlen + h] = lv[tid];
}

// ================================================================== engine
struct Engine::DPlanHolder {
  DPlan d;
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

Status Engine::create(int device, Plan&& plan, const float* regrets, const float* sums, std::unique_ptr<Engine>& out) {
  std::unique_ptr<Engine> e(new Engine());
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

  // sizes and the HBM budget (crates/solver/src/gpu/engine.rs:261-315 declines with a 512 MiB margin)
  const uint32_t rowlen = std::max(P.n[0], P.n[1]);
  e->rowlen_ = rowlen;
  size_t reach_elems = std::max((size_t)P.num_slots[0] * P.n[1], (size_t)P.num_slots[1] * P.n[0]);
  size_t slots = std::max(P.num_slots[0], P.num_slots[1]);
  size_t need = 2 * P.storage_len * sizeof(float) + reach_elems * 8 + (size_t)P.num_vslots * rowlen * 8 + slots * 4;
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  if (need + (512ull << 20) > fr) {
    char m[200];
    std::snprintf(m, sizeof m, "out of GPU memory: the solve needs %zu bytes, %zu available", need, fr);
    set_error(m);
    return kOutOfMemory;
  }

  // plan blob
  std::vector<unsigned char> blob;
  struct Off { size_t v; };
  size_t o_kind = put(blob, P.kind), o_na = put(blob, P.na), o_tfl = put(blob, P.tflags);
  size_t o_fc = put(blob, P.first_child), o_vs = put(blob, P.vslot), o_tb = put(blob, P.tboard), o_rs = put(blob, P.rslot);
  size_t o_off = put(blob, P.offset), o_inv = put(blob, P.invest), o_rake = put(blob, P.rake), o_fac = put(blob, P.factor);
  size_t o_card = put(blob, P.card), o_loc = put(blob, P.local), o_src = put(blob, P.alive_src);
  size_t o_sr = put(blob, P.seg_root), o_sp = put(blob, P.seg_parent), o_smo = put(blob, P.seg_mem_off);
  size_t o_smc = put(blob, P.seg_mem_cnt), o_mem = put(blob, P.members), o_term = put(blob, P.terminals);
  size_t o_cards[2], o_same[2], o_bo[2], o_bi[2], o_eo[2], o_ec[2], o_cb[2], o_ei[2], o_co[2], o_ci[2], o_mc[2], o_he[2], o_w[2];
  for (int q = 0; q < 2; q++) {
    o_cards[q] = put(blob, P.cards[q]);
    o_same[q] = put(blob, P.same[q]);
    o_bo[q] = put(blob, P.bl_off[q]);
    o_bi[q] = put(blob, P.bl_idx[q]);
    o_eo[q] = put(blob, P.ent_off[q]);
    o_ec[q] = put(blob, P.ent_cnt[q]);
    o_cb[q] = put(blob, P.cl_base[q]);
    o_ei[q] = put(blob, P.ent_idx[q]);
    o_co[q] = put(blob, P.cl_off[q]);
    o_ci[q] = put(blob, P.cl_idx[q]);
    o_mc[q] = put(blob, P.my_cnt[q]);
    o_he[q] = put(blob, P.hand_ent[q]);
    o_w[q] = put(blob, P.weights[q]);
  }
  DeviceBuf b;
  Status st;
  if ((st = e->alloc(b, blob.size())) != kOk) return st;
  HIPCHK(hipMemcpy(b.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
  unsigned char* base = (unsigned char*)b.p;
  e->dp_.reset(new DPlanHolder());
  DPlan& d = e->dp_->d;
  d.num_nodes = (uint32_t)P.num_nodes;
  d.n[0] = P.n[0];
  d.n[1] = P.n[1];
  d.max_entries = std::max(P.max_entries, 1u);
  d.half_pot = P.half_pot;
  d.norm = P.norm;
#define AT(T, o) ((const T*)(base + (o)))
  d.kind = AT(uint8_t, o_kind); d.na = AT(uint8_t, o_na); d.tflags = AT(uint8_t, o_tfl);
  d.first_child = AT(uint32_t, o_fc); d.vslot = AT(uint32_t, o_vs); d.tboard = AT(uint32_t, o_tb); d.rslot = AT(uint32_t, o_rs);
  d.offset = AT(uint64_t, o_off); d.invest = AT(int32_t, o_inv); d.rake = AT(float, o_rake); d.factor = AT(float, o_fac);
  d.card = AT(int8_t, o_card); d.local = AT(uint16_t, o_loc); d.alive_src = AT(uint16_t, o_src);
  d.seg_root = AT(uint32_t, o_sr); d.seg_parent = AT(uint32_t, o_sp); d.seg_mem_off = AT(uint32_t, o_smo);
  d.seg_mem_cnt = AT(uint32_t, o_smc); d.members = AT(uint32_t, o_mem); d.terminals = AT(uint32_t, o_term);
  for (int q = 0; q < 2; q++) {
    d.cards[q] = AT(uint8_t, o_cards[q]); d.same[q] = AT(uint32_t, o_same[q]);
    d.bl_off[q] = AT(uint32_t, o_bo[q]); d.bl_idx[q] = AT(uint16_t, o_bi[q]);
    d.ent_off[q] = AT(uint32_t, o_eo[q]); d.ent_cnt[q] = AT(uint32_t, o_ec[q]); d.cl_base[q] = AT(uint32_t, o_cb[q]);
    d.ent_idx[q] = AT(uint16_t, o_ei[q]); d.cl_off[q] = AT(uint16_t, o_co[q]); d.cl_idx[q] = AT(uint16_t, o_ci[q]); d.my_cnt[q] = AT(uint16_t, o_mc[q]);
    d.hand_ent[q] = AT(uint16_t, o_he[q]);
    e->weights_[q] = (float*)AT(float, o_w[q]);
  }
#undef AT
  // storage and scratch
  DeviceBuf br, bs, bre, bv, ba, bp;
  if ((st = e->alloc(br, P.storage_len * sizeof(float))) != kOk) return st;
  if ((st = e->alloc(bs, P.storage_len * sizeof(float))) != kOk) return st;
  if ((st = e->alloc(bre, reach_elems * 8)) != kOk) return st;
  if ((st = e->alloc(bv, (size_t)P.num_vslots * rowlen * 8)) != kOk) return st;
  if ((st = e->alloc(ba, slots * 4)) != kOk) return st;
  if ((st = e->alloc(bp, sizeof(IterParams) + 2 * rowlen * 8)) != kOk) return st;
  e->regrets_ = (float*)br.p;
  e->sums_ = (float*)bs.p;
  e->reach_ = bre.p;
  e->vals_ = bv.p;
  e->alive_ = (uint32_t*)ba.p;
  e->params_ = (IterParams*)bp.p;
  HIPCHK(hipHostMalloc((void**)&e->params_host_, sizeof(IterParams)));
  if ((st = e->upload(regrets, sums)) != kOk) return st;

  // LDS sizes
  e->lds_down_ = std::max<size_t>(P.max_members, 1);
  for (int f64 = 0; f64 < 2; f64++) {
    size_t vs = f64 ? 8 : 4;
    size_t mm = std::max<size_t>(P.max_entries, 1);
    e->lds_term_[f64] = (mm + std::max<size_t>(2 * mm, 53)) * 8 + (size_t)rowlen * vs;
    e->lds_up_[f64] = (size_t)P.max_members * UP_T * vs;
  }
  size_t lds_max = std::max({e->lds_term_[1], e->lds_up_[1], e->lds_down_});
  if (lds_max > 160 * 1024) {
    set_error("invalid input: a segment or strength table is too large for the LDS");
    return kInvalidInput;
  }
  if (lds_max > 64 * 1024) {
    HIPCHK(hipFuncSetAttribute((const void*)terminal_kernel<double>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    HIPCHK(hipFuncSetAttribute((const void*)terminal_kernel<float>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    HIPCHK(hipFuncSetAttribute((const void*)up_kernel<double, kBest>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    HIPCHK(hipFuncSetAttribute((const void*)up_kernel<double, kAverage>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    HIPCHK(hipFuncSetAttribute((const void*)up_kernel<float, kAverage>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    HIPCHK(hipFuncSetAttribute((const void*)up_kernel<float, kCfr>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  }
  const char* ng = std::getenv("PTS_NO_GRAPH");
  e->use_graph_ = !(ng && ng[0] == '1');
  if (e->use_graph_ && (st = e->build_graphs()) != kOk) return st;
  out = std::move(e);
  return kOk;
}

Status Engine::upload(const float* regrets, const float* sums) {
  HIPCHK(hipMemcpy(regrets_, regrets, plan_.storage_len * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(sums_, sums, plan_.storage_len * sizeof(float), hipMemcpyHostToDevice));
  return kOk;
}

Status Engine::download(float* regrets, float* sums) const {
  HIPCHK(hipStreamSynchronize(stream_));
  HIPCHK(hipMemcpy(regrets, regrets_, plan_.storage_len * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(sums, sums_, plan_.storage_len * sizeof(float), hipMemcpyDeviceToHost));
  return kOk;
}

Status Engine::read_sums(uint64_t offset, size_t len, float* out) const {
  if (offset + len > plan_.storage_len) {
    set_error("invalid input: storage range");
    return kInvalidInput;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  HIPCHK(hipMemcpy(out, sums_ + offset, len * sizeof(float), hipMemcpyDeviceToHost));
  return kOk;
}

// The root-value row of traverser p is kept past the other traversal.
static inline size_t result_offset(int p, uint32_t rowlen) { return sizeof(IterParams) + (size_t)p * rowlen * 8; }

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

template <typename V, int DMODE, int UMODE>
static void launch_span(const DPlan& d, const Plan& P, hipStream_t st, int p, Span span, uint32_t first_depth, void* reach,
                        uint32_t* alive, float* regrets, float* sums, const IterParams* prm, void* vals, uint32_t rowlen,
                        size_t lds_down, size_t lds_term, size_t lds_up, V* qout, ProfRec* prof = nullptr) {
  const uint32_t D = P.depths();
  for (uint32_t dd = first_depth; dd < D; dd++) {
    uint32_t a = P.depth_off[dd], b = P.depth_off[dd + 1];
    if (b > a) {
      PROF_START(kKDown);
      hipLaunchKernelGGL((down_kernel<V, DMODE>), dim3(b - a), dim3(DOWN_T), lds_down, st, d, p, a, span, (V*)reach, alive,
                         regrets, sums, prm);
      PROF_STOP();
    }
  }
  if (!P.terminals.empty()) {
    PROF_START(kKTerminal);
    hipLaunchKernelGGL((terminal_kernel<V>), dim3((uint32_t)P.terminals.size()), dim3(TERM_T), lds_term, st, d, p, span,
                       (const V*)reach, (const uint32_t*)alive, (V*)vals, rowlen);
    PROF_STOP();
  }
  const uint32_t tiles = (P.n[p] + UP_T - 1) / UP_T;
  for (uint32_t dd = D; dd-- > first_depth;) {
    uint32_t a = P.depth_off[dd], b = P.depth_off[dd + 1];
    if (b > a) {
      PROF_START(kKUp);
      hipLaunchKernelGGL((up_kernel<V, UMODE>), dim3(b - a, tiles), dim3(UP_T), lds_up, st, d, p, a, span, (V*)vals, rowlen,
                         (const uint32_t*)alive, regrets, (const float*)sums, prm, qout);
      PROF_STOP();
    }
  }
}

Status Engine::launch_traversal(int p, Mode mode, bool f64, ProfRec* prof) {
  const Plan& P = plan_;
  const DPlan& d = dp_->d;
  const int o = p ^ 1;
  Span span{0, 1, (uint32_t)P.num_nodes, 0};
  const uint32_t nq = P.n[o];
  PROF_START(kKBegin);
  if (f64) hipLaunchKernelGGL((begin_kernel<double>), dim3((nq + 255) / 256), dim3(256), 0, stream_, nq, weights_[o], (double*)reach_, 0u, alive_, 1u);
  else hipLaunchKernelGGL((begin_kernel<float>), dim3((nq + 255) / 256), dim3(256), 0, stream_, nq, weights_[o], (float*)reach_, 0u, alive_, 1u);
  PROF_STOP();
  if (mode == kCfr)
    launch_span<float, kCfr, kCfr>(d, P, stream_, p, span, 0, reach_, alive_, regrets_, sums_, params_, vals_, rowlen_, lds_down_,
                                   lds_term_[0], lds_up_[0], (float*)nullptr, prof);
  else if (mode == kBest)
    launch_span<double, kAverage, kBest>(d, P, stream_, p, span, 0, reach_, alive_, regrets_, sums_, params_, vals_, rowlen_,
                                         lds_down_, lds_term_[1], lds_up_[1], (double*)nullptr);
  else if (f64)
    launch_span<double, kAverage, kAverage>(d, P, stream_, p, span, 0, reach_, alive_, regrets_, sums_, params_, vals_, rowlen_,
                                            lds_down_, lds_term_[1], lds_up_[1], (double*)nullptr);
  else
    launch_span<float, kAverage, kAverage>(d, P, stream_, p, span, 0, reach_, alive_, regrets_, sums_, params_, vals_, rowlen_,
                                           lds_down_, lds_term_[0], lds_up_[0], (float*)nullptr);
  // keep the root row of p
  size_t vs = f64 ? 8 : 4;
  HIPCHK(hipMemcpyAsync((char*)params_ + result_offset(p, rowlen_), (char*)vals_ + (size_t)P.vslot[0] * rowlen_ * vs,
                        (size_t)P.n[p] * vs, hipMemcpyDeviceToDevice, stream_));
  HIPCHK(hipGetLastError());
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
  size_t n = plan_.n[p];
  out.assign(n, 0.0);
  if (f64) {
    HIPCHK(hipMemcpyAsync(out.data(), (char*)params_ + result_offset(p, rowlen_), n * 8, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
  } else {
    std::vector<float> f(n);
    HIPCHK(hipMemcpyAsync(f.data(), (char*)params_ + result_offset(p, rowlen_), n * 4, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n; i++) out[i] = f[i];
  }
  return kOk;
}

Status Engine::root_values(Mode mode, double res[2]) {  // cfr.rs:482-490, in FP64
  Status st = run_graph(mode == kBest ? 1 : 2);
  if (st != kOk) return st;
  std::vector<double> rows(2 * (size_t)rowlen_);
  HIPCHK(hipMemcpyAsync(rows.data(), (char*)params_ + result_offset(0, rowlen_), 2 * (size_t)rowlen_ * 8, hipMemcpyDeviceToHost,
                        stream_));
  HIPCHK(hipStreamSynchronize(stream_));
  for (int p = 0; p < 2; p++) {
    double s = 0.0;
    for (uint32_t i = 0; i < plan_.n[p]; i++) s += rows[(size_t)p * rowlen_ + i] * (double)plan_.weights[p][i];
    res[p] = s;
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
  const uint32_t slot = P.rslot[(size_t)q * P.num_nodes + node];
  hipLaunchKernelGGL((begin_kernel<float>), dim3((nq + 255) / 256), dim3(256), 0, stream_, nq, (const float*)dreach, (float*)reach_,
                     slot, alive_, any ? 1u : 0u);
  Span span{node, P.first_child[node], P.subtree_end[node], P.seg_of[node]};
  // depths from the node's segment on
  uint32_t seg = P.seg_of[node], first_depth = 0;
  while (first_depth + 1 < P.depth_off.size() && P.depth_off[first_depth + 1] <= seg) first_depth++;
  launch_span<float, kAverage, kAverage>(dp_->d, P, stream_, q, span, first_depth, reach_, alive_, regrets_, sums_, params_, vals_,
                                         rowlen_, lds_down_, lds_term_[0], lds_up_[0], dout);
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
