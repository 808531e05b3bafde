// The counterfactual-value network of the depth-limited flop solve: an MLP
// 2706 -> 1024 -> 1024 -> 1024 -> 1024 -> 2652 (ReLU between layers), BF16
// weights and activations, FP32 accumulation, trained with a Huber loss and
// Adam. No counterpart in the reference (a value net is new here).
//
// Input (2706): the 52-card board mask, pot / (pot + stack) and stack /
// (pot + stack), then the two players' ranges normalized to sum 1 (1326 combos
// each). Output (2652): the two players' expected values per combo, in
// fractions of the pot.
//
// Every product is one hand-written GEMM kernel on gfx950 MFMA
// (v_mfma_f32_16x16x32_bf16): 64x64 output tiles per 512-thread workgroup,
// two groups of four waves on alternate K steps of 64, each wave a 32x32
// quadrant (2x2 MFMA blocks), tiles staged through double-buffered LDS (the
// next four steps' tiles in flight in registers), tiles dealt to the XCDs in
// compact blocks. Widths are padded
// (input 2752, output 2688) so every dimension is a multiple of 64; padded
// weights stay zero.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <vector>

#include "net.hpp"

namespace pts {
namespace net {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 64, TN = 64, TK = 64;
constexpr int LDS_K = 72;  // padded LDS row (bf16): 144 bytes, 16-byte aligned

enum Epi : int {
  kEpiF32 = 0,       // C (f32) = A B
  kEpiReluBf16 = 1,  // C (bf16) = relu(A B + bias)
  kEpiBiasF32 = 2,   // C (f32) = A B + bias
  kEpiMaskBf16 = 3,  // C (bf16) = (A B) * [mask > 0]
};

// The output tile of workgroup `bid` on a grid of tn_count x tm_count tiles.
// Workgroups go to the 8 XCDs in turn (bid % 8), and each XCD has its own
// L2: so the tiles are dealt so that the workgroups of one XCD cover a
// compact block of rows and columns (groups of 4 tile rows, walked column
// by column), and each XCD fetches a fraction of A and B instead of most of
// both. Grids whose tile count is not a multiple of 8 keep row-major order.
__device__ __forceinline__ void tile_of(int bid, int tn_count, int tm_count, int& tm, int& tn) {
  const int nb = tn_count * tm_count;
  int lin = bid;
  if ((nb & 7) == 0) lin = (bid & 7) * (nb >> 3) + (bid >> 3);
  constexpr int G = 4;
  const int full = tm_count / G * G;  // tile rows in whole groups of G
  if (lin < full * tn_count) {
    const int g = lin / (G * tn_count), r = lin % (G * tn_count);
    tm = g * G + r % G;
    tn = r / G;
  } else {  // the last rows (fewer than G): row-major
    const int r = lin - full * tn_count;
    tm = full + r / tn_count;
    tn = r % tn_count;
  }
}

// C[M x N] = A[M x K] B[K x N]. A is K-contiguous (A[m * lda + k]) when AK,
// else M-contiguous (A[k * lda + m]); B is given K-contiguous as Bt[n * ldb +
// k] when BK, else N-contiguous (B[k * ldb + n]). M, N and K are multiples of
// 64 (the host checks); a 1-D grid of (M / 64) (N / 64) workgroups.
//
// 512 threads: two groups of four waves share the output tile and split its
// K steps (group g takes steps i with i % 2 == g), so each SIMD runs two
// waves whose loads, LDS traffic and MFMAs interleave; group 1's partial
// tile is added into group 0's through LDS before the epilogue.
constexpr int GT = 512;
template <bool AK, bool BK, int EPI>
__global__ void __launch_bounds__(GT) gemm_kernel(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb,
                                                  void* __restrict__ C, int ldc, int M, int N, int K,
                                                  const float* __restrict__ bias, const bf16* __restrict__ mask, int ldm) {
  // per group: double-buffered A and B tiles
  __shared__ __attribute__((aligned(16))) bf16 As[2][2][TM][LDS_K];
  __shared__ __attribute__((aligned(16))) bf16 Bs[2][2][TN][LDS_K];
  const int lane = threadIdx.x & 63, wave = (threadIdx.x >> 6) & 3, grp = threadIdx.x >> 8;
  const int tid = threadIdx.x & 255;  // thread within the group
  int tm, tn;
  tile_of(blockIdx.x, N / TN, M / TM, tm, tn);
  const int m0 = tm * TM, n0 = tn * TN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // two 16-byte global loads per operand per thread per K step: chunk
  // c = tid + 256 q of the 64 x 64 tile, 8 elements each
  auto load_a = [&](int k0, int q) -> bf16x8 {
    const int c = tid + 256 * q;
    if (AK) return *(const bf16x8*)(A + (size_t)(m0 + (c >> 3)) * lda + k0 + (c & 7) * 8);
    return *(const bf16x8*)(A + (size_t)(k0 + (c >> 3)) * lda + m0 + (c & 7) * 8);
  };
  auto load_b = [&](int k0, int q) -> bf16x8 {
    const int c = tid + 256 * q;
    if (BK) return *(const bf16x8*)(B + (size_t)(n0 + (c >> 3)) * ldb + k0 + (c & 7) * 8);
    return *(const bf16x8*)(B + (size_t)(k0 + (c >> 3)) * ldb + n0 + (c & 7) * 8);
  };
  auto store_a = [&](int buf, bf16x8 v, int q) {
    const int c = tid + 256 * q;
    if (AK) {
      *(bf16x8*)&As[grp][buf][c >> 3][(c & 7) * 8] = v;
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++) As[grp][buf][(c & 7) * 8 + j][c >> 3] = v[j];
    }
  };
  auto store_b = [&](int buf, bf16x8 v, int q) {
    const int c = tid + 256 * q;
    if (BK) {
      *(bf16x8*)&Bs[grp][buf][c >> 3][(c & 7) * 8] = v;
    } else {
#pragma unroll
      for (int j = 0; j < 8; j++) Bs[grp][buf][(c & 7) * 8 + j][c >> 3] = v[j];
    }
  };

  // The group's own steps j = 0, 1, ... are K steps 2 j + grp. Their tiles
  // are loaded into register stage j % KS, KS - 1 steps ahead of their store
  // to LDS; LDS is double-buffered (step j in buffer j % 2), so step j + 1's
  // tile is stored while step j is multiplied, and one barrier per round
  // separates a buffer's reads from its next writes. Both groups run the
  // same rounds (a group without a step in a round only meets the barrier).
  // Stage and buffer indices are constants in the unrolled loop (KS is
  // even), so the stages stay in registers.
  constexpr int KS = 4;
  const int nk = K / TK;
  const int rounds = (nk + 1) / 2;
  const int mine = (nk - grp + 1) / 2;  // this group's steps
  auto kof = [&](int j) { return (2 * j + grp) * TK; };
  bf16x8 pa[KS][2], pb[KS][2];
#pragma unroll
  for (int q = 0; q < KS; q++)
    if (q < mine) {
      pa[q][0] = load_a(kof(q), 0);
      pa[q][1] = load_a(kof(q), 1);
      pb[q][0] = load_b(kof(q), 0);
      pb[q][1] = load_b(kof(q), 1);
    }
  if (mine > 0) {
    store_a(0, pa[0][0], 0);
    store_a(0, pa[0][1], 1);
    store_b(0, pb[0][0], 0);
    store_b(0, pb[0][1], 1);
    if (KS < mine) {
      pa[0][0] = load_a(kof(KS), 0);
      pa[0][1] = load_a(kof(KS), 1);
      pb[0][0] = load_b(kof(KS), 0);
      pb[0][1] = load_b(kof(KS), 1);
    }
  }
  __syncthreads();
  const int r = lane & 15, kg = (lane >> 4) * 8;
  for (int j0 = 0; j0 < rounds; j0 += KS) {
#pragma unroll
    for (int q = 0; q < KS; q++) {
      const int j = j0 + q;
      if (j < rounds) {
        if (j < mine) {
          const int cur = q & 1, nxt = cur ^ 1, st = (q + 1) % KS;
          if (j + 1 < mine) {
            store_a(nxt, pa[st][0], 0);
            store_a(nxt, pa[st][1], 1);
            store_b(nxt, pb[st][0], 0);
            store_b(nxt, pb[st][1], 1);
            if (j + 1 + KS < mine) {
              const int k = kof(j + 1 + KS);
              pa[st][0] = load_a(k, 0);
              pa[st][1] = load_a(k, 1);
              pb[st][0] = load_b(k, 0);
              pb[st][1] = load_b(k, 1);
            }
          }
#pragma unroll
          for (int s = 0; s < TK; s += 32) {
            bf16x8 fa[2], fb[2];
#pragma unroll
            for (int ii = 0; ii < 2; ii++) fa[ii] = *(const bf16x8*)&As[grp][cur][wm + ii * 16 + r][s + kg];
#pragma unroll
            for (int jj = 0; jj < 2; jj++) fb[jj] = *(const bf16x8*)&Bs[grp][cur][wn + jj * 16 + r][s + kg];
#pragma unroll
            for (int ii = 0; ii < 2; ii++)
#pragma unroll
              for (int jj = 0; jj < 2; jj++)
                acc[ii][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[ii], fb[jj], acc[ii][jj], 0, 0, 0);
          }
        }
        __syncthreads();
      }
    }
  }
  // group 1's partial tile into group 0's: D[4 * (lane / 16) + e][lane % 16]
  // of every 16x16 block, through LDS (the A tiles, free after the last barrier)
  float* red = (float*)&As[0][0][0][0];
  static_assert(sizeof(As) >= 4 * 2 * 2 * 4 * 64 * sizeof(float), "the reduction fits the A tiles");
  if (grp == 1) {
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int e = 0; e < 4; e++) red[(((wave * 2 + i) * 2 + j) * 4 + e) * 64 + lane] = acc[i][j][e];
  }
  __syncthreads();
  if (grp == 1) return;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) acc[i][j][e] += red[(((wave * 2 + i) * 2 + j) * 4 + e) * 64 + lane];
  const int col = lane & 15, rowb = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int m = m0 + wm + i * 16 + rowb + e, n = n0 + wn + j * 16 + col;
        float x = acc[i][j][e];
        if (EPI == kEpiF32) {
          ((float*)C)[(size_t)m * ldc + n] = x;
        } else if (EPI == kEpiReluBf16) {
          x += bias[n];
          ((bf16*)C)[(size_t)m * ldc + n] = (bf16)(x > 0.f ? x : 0.f);
        } else if (EPI == kEpiBiasF32) {
          ((float*)C)[(size_t)m * ldc + n] = x + bias[n];
        } else {
          const float g = (float)mask[(size_t)m * ldm + n];
          ((bf16*)C)[(size_t)m * ldc + n] = (bf16)(g > 0.f ? x : 0.f);
        }
      }
}

// x (f32, rows of IN) -> bf16 rows of IN_P, zero padded
__global__ void pack_input_kernel(const float* __restrict__ x, bf16* __restrict__ out, int rows) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * IN_P) return;
  const int r = (int)(i / IN_P), c = (int)(i % IN_P);
  out[i] = (bf16)(c < IN ? x[(size_t)r * IN + c] : 0.f);
}

// y (f32, rows of OUT_P) -> rows of OUT
__global__ void unpack_output_kernel(const float* __restrict__ y, float* __restrict__ out, int rows) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * OUT) return;
  const int r = (int)(i / OUT), c = (int)(i % OUT);
  out[i] = y[(size_t)r * OUT_P + c];
}

// Huber loss (delta 1) over masked outputs, its gradient in bf16 (zero on
// padding and masked entries), scaled by inv_count = 1 / sum(mask); one
// partial loss sum per workgroup.
__global__ void huber_kernel(const float* __restrict__ pred, const float* __restrict__ target, const float* __restrict__ mask,
                             bf16* __restrict__ grad, float inv_count, int rows, double* __restrict__ partial) {
  __shared__ double red[256];
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < (size_t)rows * OUT_P) {
    const int r = (int)(i / OUT_P), c = (int)(i % OUT_P);
    float g = 0.f;
    if (c < OUT) {
      const size_t t = (size_t)r * OUT + c;
      const float w = mask[t];
      const float d = pred[i] - target[t];
      const float ad = fabsf(d);
      l = (double)w * (ad <= 1.f ? 0.5 * (double)d * d : (double)ad - 0.5);
      g = w * (ad <= 1.f ? d : (d > 0.f ? 1.f : -1.f)) * inv_count;
    }
    grad[i] = (bf16)g;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// db[n] = sum over rows of dz[r][n]
__global__ void colsum_kernel(const bf16* __restrict__ dz, int ld, int rows, int cols, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= cols) return;
  float s = 0.f;
  for (int r = 0; r < rows; r++) s += (float)dz[(size_t)r * ld + n];
  db[n] = s;
}

// Adam on f32 master weights; refreshes the bf16 copy.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                            bf16* __restrict__ wb, size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (bf16)wi;
}

__global__ void to_bf16_kernel(const float* __restrict__ x, bf16* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (bf16)x[i];
}

#define NCHK(x)                                                                       \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess) {                                                           \
      set_error(std::string("HIP error in " #x ": ") + hipGetErrorString(e_));       \
      return kDriver;                                                                 \
    }                                                                                 \
  } while (0)

static const int kDin[kLayers] = {IN_P, HID, HID, HID, HID};
static const int kDout[kLayers] = {HID, HID, HID, HID, OUT_P};
static const int kDinU[kLayers] = {IN, HID, HID, HID, HID};     // unpadded
static const int kDoutU[kLayers] = {HID, HID, HID, HID, OUT};

size_t num_params() {
  size_t n = 0;
  for (int l = 0; l < kLayers; l++) n += (size_t)kDoutU[l] * kDinU[l] + kDoutU[l];
  return n;
}

Net::~Net() {
  for (void* p : bufs_) (void)hipFree(p);
  if (ev_[0]) (void)hipEventDestroy(ev_[0]);
  if (ev_[1]) (void)hipEventDestroy(ev_[1]);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

// The working rows of the allocation-free forward (forward_noalloc): the
// packed bf16 input, two bf16 hidden-row buffers used in turn and the padded
// f32 output, for up to kScratchRows rows. Every GPU the library runs on
// holds one copy until the process ends; all calls on a GPU share it, so
// calls that overlap on one GPU race. Host code never names the array: its
// device address is taken per device at Net::create (hipGetSymbolAddress,
// the device current) and passed to the kernels.
constexpr size_t kScrIn = (size_t)kScratchRows * IN_P * 2;
constexpr size_t kScrHid = (size_t)kScratchRows * HID * 2;
constexpr size_t kScrOut = (size_t)kScratchRows * OUT_P * 4;
static __device__ __align__(256) unsigned char scratch[kScrIn + 2 * kScrHid + kScrOut];
constexpr int kMaxDevices = 64;
static unsigned char* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// The scratch array's address on `device` (current), taken once per device.
static Status scratch_of(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("invalid input: device index beyond the scratch table");
    return kInvalidInput;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* ptr = nullptr;
    NCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[device] = (unsigned char*)ptr;
  }
  *out = g_scratch[device];
  return kOk;
}

Status Net::dalloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_error("out of GPU memory for the value network");
    return kOutOfMemory;
  }
  (void)hipMemset(*p, 0, bytes ? bytes : 1);
  bufs_.push_back(*p);
  return kOk;
}

Status Net::create(int device, uint32_t batch_cap, std::unique_ptr<Net>& out) {
  if (batch_cap == 0 || batch_cap % 64 != 0) {
    set_error("invalid input: the network's batch capacity is a positive multiple of 64");
    return kInvalidInput;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device >= count) {
    set_error("no GPU device");
    return kNoDevice;
  }
  NCHK(hipSetDevice(device));
  std::unique_ptr<Net> n(new Net());
  n->cap_ = batch_cap;
  Status sst = scratch_of(device, &n->scratch_);
  if (sst != kOk) return sst;
  NCHK(hipStreamCreateWithFlags(&n->stream_, hipStreamNonBlocking));
  n->own_stream_ = true;
  NCHK(hipEventCreate(&n->ev_[0]));
  NCHK(hipEventCreate(&n->ev_[1]));
  Status st;
  for (int l = 0; l < kLayers; l++) {
    const size_t wn = (size_t)kDout[l] * kDin[l];
    void* p;
    if ((st = n->dalloc(&p, wn * 4)) != kOk) return st;
    n->w_[l] = (float*)p;
    if ((st = n->dalloc(&p, wn * 2)) != kOk) return st;
    n->wb_[l] = (bf16*)p;
    if ((st = n->dalloc(&p, wn * 4)) != kOk) return st;
    n->gw_[l] = (float*)p;
    if ((st = n->dalloc(&p, wn * 4)) != kOk) return st;
    n->mw_[l] = (float*)p;
    if ((st = n->dalloc(&p, wn * 4)) != kOk) return st;
    n->vw_[l] = (float*)p;
    const size_t bn = kDout[l];
    if ((st = n->dalloc(&p, bn * 4)) != kOk) return st;
    n->b_[l] = (float*)p;
    if ((st = n->dalloc(&p, bn * 4)) != kOk) return st;
    n->gb_[l] = (float*)p;
    if ((st = n->dalloc(&p, bn * 4)) != kOk) return st;
    n->mb_[l] = (float*)p;
    if ((st = n->dalloc(&p, bn * 4)) != kOk) return st;
    n->vb_[l] = (float*)p;
  }
  // activations: act_[0] = packed input, act_[l] = input of layer l
  for (int l = 0; l < kLayers; l++) {
    void* p;
    if ((st = n->dalloc(&p, (size_t)batch_cap * kDin[l] * 2)) != kOk) return st;
    n->act_[l] = (bf16*)p;
    if ((st = n->dalloc(&p, (size_t)batch_cap * kDout[l] * 2)) != kOk) return st;
    n->dz_[l] = (bf16*)p;
  }
  void* p;
  if ((st = n->dalloc(&p, (size_t)batch_cap * OUT_P * 4)) != kOk) return st;
  n->y_ = (float*)p;
  if ((st = n->dalloc(&p, (size_t)batch_cap * IN * 4)) != kOk) return st;
  n->xin_ = (float*)p;
  if ((st = n->dalloc(&p, (size_t)batch_cap * OUT * 4 * 3)) != kOk) return st;
  n->yout_ = (float*)p;
  n->tgt_ = n->yout_ + (size_t)batch_cap * OUT;
  n->msk_ = n->tgt_ + (size_t)batch_cap * OUT;
  const size_t parts = ((size_t)batch_cap * OUT_P + 255) / 256;
  if ((st = n->dalloc(&p, parts * 8)) != kOk) return st;
  n->partial_ = (double*)p;
  out = std::move(n);
  return kOk;
}

// packed host layout: W0 (1024 x 2706, row-major), b0, W1, b1, ..., W4 (2652 x 1024), b4
Status Net::set_params(const float* params) {
  const float* q = params;
  for (int l = 0; l < kLayers; l++) {
    std::vector<float> w((size_t)kDout[l] * kDin[l], 0.f), b(kDout[l], 0.f);
    for (int r = 0; r < kDoutU[l]; r++)
      for (int c = 0; c < kDinU[l]; c++) w[(size_t)r * kDin[l] + c] = *q++;
    for (int r = 0; r < kDoutU[l]; r++) b[r] = *q++;
    NCHK(hipMemcpy(w_[l], w.data(), w.size() * 4, hipMemcpyHostToDevice));
    NCHK(hipMemcpy(b_[l], b.data(), b.size() * 4, hipMemcpyHostToDevice));
    const size_t wn = w.size();
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((wn + 255) / 256)), dim3(256), 0, stream_, (const float*)w_[l], wb_[l], wn);
    NCHK(hipMemsetAsync(mw_[l], 0, wn * 4, stream_));
    NCHK(hipMemsetAsync(vw_[l], 0, wn * 4, stream_));
    NCHK(hipMemsetAsync(mb_[l], 0, b.size() * 4, stream_));
    NCHK(hipMemsetAsync(vb_[l], 0, b.size() * 4, stream_));
  }
  step_ = 0;
  NCHK(hipStreamSynchronize(stream_));
  return kOk;
}

static Status unpack(const float* const* dev, const int* rows, const int* cols, const int* ld, int count, float* host,
                     bool weights_then_bias, const float* const* bdev) {
  (void)weights_then_bias;
  float* q = host;
  for (int l = 0; l < count; l++) {
    std::vector<float> w((size_t)rows[l] * ld[l]);
    NCHK(hipMemcpy(w.data(), dev[l], w.size() * 4, hipMemcpyDeviceToHost));
    for (int r = 0; r < kDoutU[l]; r++)
      for (int c = 0; c < kDinU[l]; c++) *q++ = w[(size_t)r * ld[l] + c];
    std::vector<float> b(rows[l]);
    NCHK(hipMemcpy(b.data(), bdev[l], b.size() * 4, hipMemcpyDeviceToHost));
    for (int r = 0; r < kDoutU[l]; r++) *q++ = b[r];
    (void)cols;
  }
  return kOk;
}

Status Net::get_params(float* params) const {
  NCHK(hipStreamSynchronize(stream_));
  return unpack(w_, kDout, kDin, kDin, kLayers, params, true, b_);
}

Status Net::get_grads(float* grads) const {
  NCHK(hipStreamSynchronize(stream_));
  return unpack(gw_, kDout, kDin, kDin, kLayers, grads, true, gb_);
}

template <bool AK, bool BK, int EPI>
static void gemm(hipStream_t s, const bf16* A, int lda, const bf16* B, int ldb, void* C, int ldc, int M, int N, int K,
                 const float* bias = nullptr, const bf16* mask = nullptr, int ldm = 0) {
  hipLaunchKernelGGL((gemm_kernel<AK, BK, EPI>), dim3((N / TN) * (M / TM)), dim3(GT), 0, s, A, lda, B, ldb, C, ldc, M, N, K,
                     bias, mask, ldm);
}

Status Net::check_rows(uint32_t rows) const {
  if (rows == 0 || rows % 64 != 0 || rows > cap_) {
    set_error("invalid input: network batch rows must be a positive multiple of 64 within the capacity");
    return kInvalidInput;
  }
  return kOk;
}

// forward on the packed input act_[0]; outputs y_ (f32, rows of OUT_P)
void Net::forward_packed(uint32_t rows) {
  const int B = (int)rows;
  for (int l = 0; l < kLayers; l++) {
    if (l + 1 < kLayers)
      gemm<true, true, kEpiReluBf16>(stream_, act_[l], kDin[l], wb_[l], kDin[l], act_[l + 1], kDout[l], B, kDout[l], kDin[l], b_[l]);
    else
      gemm<true, true, kEpiBiasF32>(stream_, act_[l], kDin[l], wb_[l], kDin[l], y_, kDout[l], B, kDout[l], kDin[l], b_[l]);
  }
}

Status Net::forward_device(const bf16* x, uint32_t rows, float* y, hipStream_t s) {
  Status st = check_rows(rows);
  if (st != kOk) return st;
  hipStream_t keep = stream_;
  stream_ = s;
  NCHK(hipMemcpyAsync(act_[0], x, (size_t)rows * IN_P * 2, hipMemcpyDeviceToDevice, s));
  forward_packed(rows);
  hipLaunchKernelGGL(unpack_output_kernel, dim3((unsigned)(((size_t)rows * OUT + 255) / 256)), dim3(256), 0, s, (const float*)y_, y,
                     (int)rows);
  stream_ = keep;
  NCHK(hipGetLastError());
  return kOk;
}

Status Net::forward_noalloc(const float* x, uint32_t rows, float* y, hipStream_t s) {
  if (rows == 0 || rows % 64 != 0 || rows > kScratchRows) {
    set_error("invalid input: the allocation-free forward takes a positive multiple of 64 rows, at most " +
              std::to_string(kScratchRows));
    return kInvalidInput;
  }
  if (!x || !y) {
    set_error("invalid input: null rows");
    return kInvalidInput;
  }
  if (!s) s = stream_;
  bf16* in = (bf16*)scratch_;
  bf16* hid[2] = {(bf16*)(scratch_ + kScrIn), (bf16*)(scratch_ + kScrIn + kScrHid)};
  float* yf = (float*)(scratch_ + kScrIn + 2 * kScrHid);
  const int B = (int)rows;
  hipLaunchKernelGGL(pack_input_kernel, dim3((unsigned)(((size_t)rows * IN_P + 255) / 256)), dim3(256), 0, s, x, in, B);
  const bf16* a = in;
  for (int l = 0; l < kLayers; l++) {
    if (l + 1 < kLayers) {
      gemm<true, true, kEpiReluBf16>(s, a, kDin[l], wb_[l], kDin[l], hid[l & 1], kDout[l], B, kDout[l], kDin[l], b_[l]);
      a = hid[l & 1];
    } else {
      gemm<true, true, kEpiBiasF32>(s, a, kDin[l], wb_[l], kDin[l], yf, kDout[l], B, kDout[l], kDin[l], b_[l]);
    }
  }
  hipLaunchKernelGGL(unpack_output_kernel, dim3((unsigned)(((size_t)rows * OUT + 255) / 256)), dim3(256), 0, s, (const float*)yf, y,
                     B);
  NCHK(hipGetLastError());
  return kOk;
}

Status Net::forward(const float* x, uint32_t rows, float* y) {
  Status st = check_rows(rows);
  if (st != kOk) return st;
  NCHK(hipMemcpyAsync(xin_, x, (size_t)rows * IN * 4, hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(pack_input_kernel, dim3((unsigned)(((size_t)rows * IN_P + 255) / 256)), dim3(256), 0, stream_,
                     (const float*)xin_, act_[0], (int)rows);
  forward_packed(rows);
  hipLaunchKernelGGL(unpack_output_kernel, dim3((unsigned)(((size_t)rows * OUT + 255) / 256)), dim3(256), 0, stream_,
                     (const float*)y_, yout_, (int)rows);
  NCHK(hipMemcpyAsync(y, yout_, (size_t)rows * OUT * 4, hipMemcpyDeviceToHost, stream_));
  NCHK(hipStreamSynchronize(stream_));
  return kOk;
}

// loss and gradients of the current batch (inputs already packed)
Status Net::backward_packed(uint32_t rows, float inv_count, double* loss) {
  const int B = (int)rows;
  forward_packed(rows);
  const size_t parts = ((size_t)rows * OUT_P + 255) / 256;
  hipLaunchKernelGGL(huber_kernel, dim3((unsigned)parts), dim3(256), 0, stream_, (const float*)y_, (const float*)tgt_,
                     (const float*)msk_, dz_[kLayers - 1], inv_count, (int)rows, partial_);
  for (int l = kLayers - 1; l >= 0; l--) {
    // dW_l = dZ_l^T act_l (f32); db_l = column sums of dZ_l
    gemm<false, false, kEpiF32>(stream_, dz_[l], kDout[l], act_[l], kDin[l], gw_[l], kDin[l], kDout[l], kDin[l], B);
    hipLaunchKernelGGL(colsum_kernel, dim3((kDout[l] + 255) / 256), dim3(256), 0, stream_, (const bf16*)dz_[l], kDout[l], B,
                       kDout[l], gb_[l]);
    // dZ_{l-1} = (dZ_l W_l) * [act_l > 0]
    if (l > 0)
      gemm<true, false, kEpiMaskBf16>(stream_, dz_[l], kDout[l], wb_[l], kDin[l], dz_[l - 1], kDin[l], B, kDin[l], kDout[l], nullptr,
                                      act_[l], kDin[l]);
  }
  std::vector<double> part(parts);
  NCHK(hipMemcpyAsync(part.data(), partial_, parts * 8, hipMemcpyDeviceToHost, stream_));
  NCHK(hipStreamSynchronize(stream_));
  double s = 0.0;
  for (double x : part) s += x;
  if (loss) *loss = s * (double)inv_count;
  return kOk;
}

Status Net::stage_batch(const float* x, const float* target, const float* mask, uint32_t rows, float* inv_count) {
  Status st = check_rows(rows);
  if (st != kOk) return st;
  double cnt = 0.0;
  for (size_t i = 0; i < (size_t)rows * OUT; i++) cnt += mask[i];
  if (!(cnt > 0.0)) {
    set_error("invalid input: the loss mask is empty");
    return kInvalidInput;
  }
  *inv_count = (float)(1.0 / cnt);
  NCHK(hipMemcpyAsync(xin_, x, (size_t)rows * IN * 4, hipMemcpyHostToDevice, stream_));
  NCHK(hipMemcpyAsync(tgt_, target, (size_t)rows * OUT * 4, hipMemcpyHostToDevice, stream_));
  NCHK(hipMemcpyAsync(msk_, mask, (size_t)rows * OUT * 4, hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(pack_input_kernel, dim3((unsigned)(((size_t)rows * IN_P + 255) / 256)), dim3(256), 0, stream_,
                     (const float*)xin_, act_[0], (int)rows);
  return kOk;
}

Status Net::gradients(const float* x, const float* target, const float* mask, uint32_t rows, double* loss) {
  float inv = 0.f;
  Status st = stage_batch(x, target, mask, rows, &inv);
  if (st != kOk) return st;
  return backward_packed(rows, inv, loss);
}

Status Net::adam(float lr) {
  step_ += 1;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
  for (int l = 0; l < kLayers; l++) {
    const size_t wn = (size_t)kDout[l] * kDin[l];
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((wn + 255) / 256)), dim3(256), 0, stream_, w_[l], (const float*)gw_[l], mw_[l],
                       vw_[l], wb_[l], wn, lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(adam_kernel, dim3((kDout[l] + 255) / 256), dim3(256), 0, stream_, b_[l], (const float*)gb_[l], mb_[l], vb_[l],
                       (bf16*)nullptr, (size_t)kDout[l], lr, b1, b2, eps, c1, c2);
  }
  NCHK(hipGetLastError());
  return kOk;
}

Status Net::train_step(const float* x, const float* target, const float* mask, uint32_t rows, float lr, double* loss) {
  float inv = 0.f;
  Status st = stage_batch(x, target, mask, rows, &inv);
  if (st != kOk) return st;
  if ((st = backward_packed(rows, inv, loss)) != kOk) return st;
  return adam(lr);
}

// FLOPs of one forward pass per row (2 per multiply-add, padded widths)
double forward_flops_per_row() {
  double f = 0.0;
  for (int l = 0; l < kLayers; l++) f += 2.0 * kDin[l] * kDout[l];
  return f;
}

Status Net::profile(uint32_t rows, double* fwd_ms, double* train_ms) {
  Status st = check_rows(rows);
  if (st != kOk) return st;
  // time the forward and one forward + backward (+ Adam on zero-step state
  // is skipped: the weights stay as they are) on the resident batch
  float ms = 0.f;
  forward_packed(rows);
  NCHK(hipEventRecord(ev_[0], stream_));
  for (int i = 0; i < 5; i++) forward_packed(rows);
  NCHK(hipEventRecord(ev_[1], stream_));
  NCHK(hipEventSynchronize(ev_[1]));
  NCHK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  *fwd_ms = ms / 5.0;
  const int B = (int)rows;
  NCHK(hipEventRecord(ev_[0], stream_));
  for (int i = 0; i < 5; i++) {
    forward_packed(rows);
    for (int l = kLayers - 1; l >= 0; l--) {
      gemm<false, false, kEpiF32>(stream_, dz_[l], kDout[l], act_[l], kDin[l], gw_[l], kDin[l], kDout[l], kDin[l], B);
      if (l > 0)
        gemm<true, false, kEpiMaskBf16>(stream_, dz_[l], kDout[l], wb_[l], kDin[l], dz_[l - 1], kDin[l], B, kDin[l], kDout[l],
                                        nullptr, act_[l], kDin[l]);
    }
  }
  NCHK(hipEventRecord(ev_[1], stream_));
  NCHK(hipEventSynchronize(ev_[1]));
  NCHK(hipEventElapsedTime(&ms, ev_[0], ev_[1]));
  *train_ms = ms / 5.0;
  return kOk;
}

void he_init(uint64_t seed, std::vector<float>& params) {
  params.assign(num_params(), 0.f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.f, 1.f);
  size_t q = 0;
  for (int l = 0; l < kLayers; l++) {
    const float sd = std::sqrt(2.f / (float)kDinU[l]);
    for (size_t i = 0; i < (size_t)kDoutU[l] * kDinU[l]; i++) params[q++] = nd(rng) * sd;
    for (int i = 0; i < kDoutU[l]; i++) params[q++] = 0.f;
  }
}

}  // namespace net
}  // namespace pts
