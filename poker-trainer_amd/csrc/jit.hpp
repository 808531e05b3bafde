// Walks specialised to a segment shape, compiled at engine creation.
//
// The generic up walk (up_reg_kernel) reads a segment's member records at
// run time and keeps the member values in register vectors indexed by a
// wave-uniform member number (indexed register moves, a few scalar
// instructions and a branch per access). A tree such as the flop spot's has
// hundreds of thousands of river segments but only a handful of distinct
// shapes (the same betting tree under each river card), so the walk of each
// frequent shape is generated as straight-line HIP with every member a named
// value, compiled once with hiprtc for gfx950, and launched over the list of
// segments of that shape. The arithmetic is the generic walk's, operation for
// operation (the same compiler flags: no contraction, no packed FP32), so
// results are bit-identical to it; only the addressing is resolved at
// compile time.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "engine.hpp"

namespace pts {
namespace jit {

// One member of a segment shape: a decision node (kind 0 or 1, its action
// count and the local index of its first child), or a member whose value is
// read from its row (kind 2: a chance node, terminal or leaf).
struct Member {
  uint8_t kind, na;
  uint16_t cl;
  bool operator==(const Member& o) const { return kind == o.kind && na == o.na && cl == o.cl; }
};
using Shape = std::vector<Member>;

// The kernels' one argument (the generated source declares the same layout).
struct WalkArgs {
  const uint32_t* list;        // segment ids list[i0 .. i0 + nseg)
  const uint4* seg_head;       // per segment: first member, member count, root node, hands of both players
  const unsigned char* mrec;   // MemRec array (48 B each)
  const uint8_t* malive;       // per member: alive for the current traverser
  float* vals;                 // value rows (f64 for the best response)
  float* regrets;              // f32 regret storage (DCFR walk)
  const float* prm;            // IterParams: pos, neg (DCFR walk)
  const float* zrow;           // 8 * rowlen f32 zeros: what a dead member reads
  int16_t* rq;                 // 16-bit regret storage (16-bit walks)
  const float* rsc;            // per player node: the regrets' scale (16-bit walks)
  int32_t* rnx;                // per player node: the new scale's float bits (16-bit walks)
  uint32_t i0, nseg, tiles;    // the segments walked; 64-hand tiles per segment
  uint32_t rowlen, v0, vb;     // row length, value-row bias (Bias::v)
};

// The kernels of one shape for one traverser: the f32 DCFR walk, the FP64
// best response, and with 16-bit storage the DCFR walk's two passes (the
// largest new regret of every own node, then the new regrets encoded with it).
enum WalkKind { kWalkF32 = 0, kWalkBest = 1, kWalkMax16 = 2, kWalkEnc16 = 3 };
constexpr int kWalkKinds = 4;
inline size_t walk_fn(size_t shape, int kind, int p) { return (shape * kWalkKinds + (size_t)kind) * 2 + (size_t)p; }

// The most members a specialised walk takes.
constexpr uint32_t kMaxMembers = 80;

// Launch grid of a walk over n segments of t hand tiles: one 64-thread
// workgroup per (segment, tile), dealt to the 8 XCDs in turn, padded to a
// multiple of 8 segments.
inline uint32_t walk_groups(uint32_t n, uint32_t t) { return (n + 7u) / 8u * 8u * t; }

// Compiles the walks of `shapes` for both traversers on the current device
// (or takes them from this process's cache): fn[walk_fn(i, kind, p)] is shape
// i's kernel of that kind for traverser p. With `q16` the kinds are the best
// response and the two 16-bit passes, else the f32 walk and the best
// response; the others are null. `seconds` gets the compile time (0 when
// cached).
Status compile_walks(const std::vector<Shape>& shapes, bool q16, std::vector<hipFunction_t>& fn, double* seconds);

// The generated HIP source for `shapes` (for inspection and tests).
std::string walk_source(const std::vector<Shape>& shapes, bool q16);

}  // namespace jit
}  // namespace pts
