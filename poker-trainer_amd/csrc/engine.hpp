// The HIP engine: plan and storage resident in HBM, one DCFR iteration per
// hipGraph launch, FP64 best response. Replaces the reference's CUDA
// `gpu::Engine` (crates/solver/src/gpu/engine.rs:291-749).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "plan.hpp"

namespace pts {
namespace net {
class Net;
}
struct Store;

enum Status : int {
  kOk = 0,
  kDriverMissing = 1,
  kDriverTooOld = 2,
  kNoDevice = 3,
  kDriver = 4,
  kInvalidInput = 5,
  kOutOfMemory = 6,
};

void set_error(const std::string& msg);
const std::string& last_error();
// The last failure's numbers (thread-local, cleared by set_error): the HIP
// error code of a driver failure, and the bytes needed / available of an
// out-of-memory failure (crates/solver/src/gpu/mod.rs:86-136).
struct ErrorDetail {
  int32_t hip_error = 0;
  uint64_t needed = 0, available = 0;
};
void set_error_detail(const ErrorDetail& d);
const ErrorDetail& last_error_detail();

// Traversal modes (crates/solver/src/cfr.rs:248-256)
enum Mode : int { kCfr = 0, kBest = 1, kAverage = 2 };

// Terminal-sum precision (crates/solver/src/gpu/mod.rs:51-63)
enum Precision : int { kFast = 0, kExact = 1 };

struct IterParams {
  float pos, neg, strat;
  int reach_weighted;
};

// Kernel classes timed by Engine::profile_iteration
enum KernelClass : int { kKBegin = 0, kKDown = 1, kKTerminal = 2, kKUp = 3, kKClasses = 4 };
struct Profile {
  double ms[kKClasses] = {0, 0, 0, 0};      // summed event time per class
  uint32_t launches[kKClasses] = {0, 0, 0, 0};
  double wall_ms = 0;                       // first event to last
};

struct DeviceBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

class Engine {
 public:
  ~Engine();
  // Uploads the plan and the storage; fails with kOutOfMemory when HBM is short.
  // precision: kFast (block-scan terminal sums, f64) or kExact (the
  // reference's sequential order, bit-identical to the CPU solve).
  // regrets / sums may be null (zeros). scratch_budget caps the reach and
  // value rows in bytes (0: all free HBM after the storage, less a margin);
  // the chunks below the root street's chance nodes run in groups that fit.
  // q16: 16-bit storage (crates/solver/src/storage.rs:87-106: i16 regrets,
  // u16 sums, one f32 scale per player node; the plan must carry player
  // slots), starting from zeros (regrets / sums must be null).
  static Status create(int device, Plan&& plan, const float* regrets, const float* sums, int precision,
                       uint64_t scratch_budget, std::unique_ptr<Engine>& out, bool q16 = false);
  int precision() const { return precision_; }
  bool compressed() const { return q16_; }

  // One DCFR iteration: player 0's update, then player 1's (crates/solver/src/lib.rs:577-585).
  Status cfr_iteration(const IterParams& d);
  // The same iteration launched kernel by kernel with HIP events around every
  // launch (on the engine's stream): per-class durations for roofline figures.
  Status profile_iteration(const IterParams& d, Profile& out);
  // One traversal of player p; CFR mode updates storage. Root values of p's hands into out (n_p).
  Status traverse(int p, Mode mode, const IterParams& d, std::vector<double>& out);
  // Expected payoffs (chips) at the root under `mode` (kBest / kAverage), traversals in FP64.
  Status root_values(Mode mode, double out[2]);
  // Every spot's per-hand root values (FP64 traversals): rows[(spot * 2 + p) * rowlen + hand].
  Status root_rows(Mode mode, std::vector<double>& rows);
  uint32_t rowlen() const { return rowlen_; }
  // Every spot's root values (a batch): out[spot * 2 + player].
  Status root_values_all(Mode mode, std::vector<double>& out);
  // Spots with active = 0 are skipped by every traversal (their roots are dead).
  Status set_active(const std::vector<uint32_t>& active);
  // Freezes (or thaws) decision nodes: a frozen node's regrets are not
  // updated by CFR traversals (its strategy stays what its regrets give).
  Status set_frozen(const std::vector<uint32_t>& nodes, bool on);
  // Tiles: terminals per fast-terminal workgroup (run length, 1..16) and
  // opponent hands per down-kernel workgroup (128, 256, 512 or 1024); rebuilds
  // the graphs. Same scratch groups (their rows do not depend on the runs).
  Status retune(uint32_t run_length, uint32_t down_threads);
  uint32_t run_length() const { return run_length_; }
  uint32_t down_threads() const { return down_threads_; }
  // A storage range to the host (either pointer may be null); f32 storage.
  Status read_storage(uint64_t offset, size_t len, float* regrets, float* sums) const;
  // 16-bit storage: the stored integers and the per-player-node scales.
  Status download_q(int16_t* regrets, uint16_t* sums, float* regret_scale, float* strat_scale) const;
  Status upload_q(const int16_t* regrets, const uint16_t* sums, const float* regret_scale, const float* strat_scale);
  Status zero_storage();
  Status write_storage(uint64_t offset, size_t len, const float* regrets, const float* sums);
  // Depth-limited trees: the value network that values the leaves (null:
  // zeros), with the root board; rebuilds the graphs. The network must stay
  // alive while attached.
  Status attach_net(net::Net* n, const std::vector<int>& board);
  Status exploitability(bool raked, double* out);
  Status download(float* regrets, float* sums) const;
  Status upload(const float* regrets, const float* sums);
  // A slice of the strategy sums (e.g. one node's block) to the host; with
  // 16-bit storage decoded with player node pslot's scale (storage.rs:236-246).
  Status read_sums(uint64_t offset, size_t len, float* out, uint32_t pslot = 0xFFFFFFFFu) const;
  // Subtree traversal for queries: player q at decision node `node`, opponent
  // reach given; writes the node's children values (na * n_q, action-major).
  Status subtree_values(uint32_t node, int q, const std::vector<float>& opp_reach, std::vector<float>& child_vals);

  const Plan& plan() const { return plan_; }
  hipStream_t stream() const { return stream_; }
  // Event-timed kernel durations of the last graph-less CFR iteration (profiling aid)
  double last_ms = 0;
  size_t device_bytes() const { return bytes_; }
  // Chunk groups of a traversal, for f32 (CFR) or f64 rows.
  const std::vector<ChunkGroup>& groups(bool f64) const { return groups_[f64 ? 1 : 0]; }
  // Segments whose CFR walks run in shape-specialised kernels, their shapes,
  // and the seconds their compilation took at creation (0 when cached).
  uint64_t jit_segments() const { return jit_segments_; }
  uint32_t jit_shapes() const { return jit_shapes_; }
  double jit_seconds() const { return jit_seconds_; }

 private:
  Engine() = default;
  Status alloc(DeviceBuf& b, size_t bytes);

  Status launch_traversal(int p, Mode mode, bool f64, struct ProfRec* prof = nullptr);
  Status setup_jit(bool q16);
  void launch_leaves(int p, Mode mode, bool f64, void* reach, void* vals, uint32_t* alive);
  Status build_graphs();
  Status run_graph(int which);

  Plan plan_;
  int device_ = 0;
  hipStream_t stream_ = nullptr;
  std::vector<DeviceBuf> bufs_;
  size_t bytes_ = 0;
  // device pointers
  void* dplan_blob_ = nullptr;
  float* regrets_ = nullptr;
  float* sums_ = nullptr;
  // 16-bit storage
  bool q16_ = false;
  int16_t* rq_ = nullptr;
  uint16_t* sq_ = nullptr;
  float *rsc_ = nullptr, *ssc_ = nullptr;
  int32_t* rnx_ = nullptr;
  Store store() const;
  // scratch pool: reach rows, value rows and alive flags for f32 [0] and f64 [1] rows
  void* reach_[2] = {nullptr, nullptr};
  void* vals_[2] = {nullptr, nullptr};
  uint32_t* alive_[2] = {nullptr, nullptr};
  size_t pool_bytes_[2] = {0, 0};
  uint32_t* active_ = nullptr;  // device, one flag per spot
  TermBatch* tbatch_dev_ = nullptr;
  uint32_t down_threads_ = 512, run_length_ = 1;
  uint64_t budget_ = 0;
  // depth-limited leaves
  net::Net* net_ = nullptr;
  const LeafRec* leaves_ = nullptr;
  int8_t *deal_ = nullptr, *board_ = nullptr;
  void *own_ = nullptr, *mreach_ = nullptr;
  __bf16* net_in_ = nullptr;
  float* net_out_ = nullptr;
  uint32_t ncards_ = 0, board_len_ = 0, net_rows_ = 0;
  IterParams* params_ = nullptr;
  IterParams* params_host_ = nullptr;
  uint32_t rowlen_ = 0;
  size_t lds_down_ = 0, lds_term_[2] = {0, 0}, lds_up_[2] = {0, 0};
  hipGraphExec_t graph_[3] = {nullptr, nullptr, nullptr};  // cfr iteration, best response, average
  bool use_graph_ = true;
  int precision_ = kFast;
  std::vector<ChunkGroup> groups_[2];
  uint64_t jit_segments_ = 0;
  uint32_t jit_shapes_ = 0;
  double jit_seconds_ = 0;
  struct DPlanHolder;
  std::unique_ptr<DPlanHolder> dp_;
};

}  // namespace pts
