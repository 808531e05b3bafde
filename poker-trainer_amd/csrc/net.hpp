// The counterfactual-value network (see net.hip): device buffers, training
// and inference entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "engine.hpp"

namespace pts {
namespace net {

constexpr int IN = 2706, IN_P = 2752;   // 52 board + 2 pot/stack + 2 x 1326 ranges; padded
constexpr int HID = 1024;
constexpr int OUT = 2652, OUT_P = 2688;  // 2 x 1326 values; padded
constexpr int kLayers = 5;
constexpr uint32_t kScratchRows = 1024;  // rows of the allocation-free forward

size_t num_params();
double forward_flops_per_row();
// He-normal weights, zero biases, in the packed layout of Net::set_params.
void he_init(uint64_t seed, std::vector<float>& params);

class Net {
 public:
  ~Net();
  // Buffers for batches of up to batch_cap rows (a multiple of 64).
  static Status create(int device, uint32_t batch_cap, std::unique_ptr<Net>& out);
  // Packed f32 parameters: W0 (1024 x 2706, row-major), b0, W1, b1, W2, b2,
  // W3, b3, W4 (2652 x 1024), b4. Resets the Adam state.
  Status set_params(const float* params);
  Status get_params(float* params) const;
  // Gradients of the last gradients() / train_step(), packed like the parameters.
  Status get_grads(float* grads) const;
  // Host rows (rows x 2706) -> host rows (rows x 2652); rows a multiple of 64.
  Status forward(const float* x, uint32_t rows, float* y);
  // Device rows (f32, rows x IN) -> device rows (f32, rows x OUT) on stream s
  // (null: the network's), with the working rows in the per-GPU scratch
  // array instead of the network's buffers: no allocation, no
  // synchronization; rows a multiple of 64 up to kScratchRows.
  Status forward_noalloc(const float* x, uint32_t rows, float* y, hipStream_t s);
  // Device rows already packed (bf16, rows x IN_P) -> device rows (f32, rows x OUT), on stream s.
  Status forward_device(const __bf16* x, uint32_t rows, float* y, hipStream_t s);
  // Masked Huber loss and its gradients for one batch (no update).
  Status gradients(const float* x, const float* target, const float* mask, uint32_t rows, double* loss);
  // gradients() then one Adam step (beta 0.9 / 0.999, eps 1e-8).
  Status train_step(const float* x, const float* target, const float* mask, uint32_t rows, float lr, double* loss);
  // Event-timed forward and forward + backward passes on the resident batch.
  Status profile(uint32_t rows, double* fwd_ms, double* train_ms);
  hipStream_t stream() const { return stream_; }
  uint32_t capacity() const { return cap_; }

 private:
  Net() = default;
  Status dalloc(void** p, size_t bytes);
  Status check_rows(uint32_t rows) const;
  Status stage_batch(const float* x, const float* target, const float* mask, uint32_t rows, float* inv_count);
  void forward_packed(uint32_t rows);
  Status backward_packed(uint32_t rows, float inv_count, double* loss);
  Status adam(float lr);

  uint32_t cap_ = 0;
  unsigned char* scratch_ = nullptr;  // this device's scratch array (device address)
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  std::vector<void*> bufs_;
  float *w_[kLayers] = {}, *b_[kLayers] = {}, *gw_[kLayers] = {}, *gb_[kLayers] = {};
  float *mw_[kLayers] = {}, *vw_[kLayers] = {}, *mb_[kLayers] = {}, *vb_[kLayers] = {};
  __bf16* wb_[kLayers] = {};
  __bf16* act_[kLayers] = {};  // act_[l]: the input rows of layer l (act_[0]: the packed input)
  __bf16* dz_[kLayers] = {};   // gradient of layer l's output (pre-activation)
  float* y_ = nullptr;         // output rows (f32, OUT_P)
  float *xin_ = nullptr, *yout_ = nullptr, *tgt_ = nullptr, *msk_ = nullptr;
  double* partial_ = nullptr;
  uint32_t step_ = 0;
};

}  // namespace net
}  // namespace pts
