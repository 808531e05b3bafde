// Shape-specialised CFR up walks: HIP source generated per segment shape and
// compiled with hiprtc (see jit.hpp).
#include "jit.hpp"

#include <hip/hiprtc.h>

#include <chrono>
#include <cstddef>
#include <map>
#include <mutex>
#include <sstream>

namespace pts {
namespace jit {

static_assert(offsetof(MemRec, offset) == 0 && offsetof(MemRec, vslot) == 12 && offsetof(MemRec, tflags) == 41 &&
                  offsetof(MemRec, pslot) == 44 && sizeof(MemRec) == 48,
              "the generated walks read MemRec at these byte offsets");
static_assert(kNodeFrozen == 8, "the generated walks test bit 3 of tflags");
static_assert(offsetof(IterParams, pos) == 0 && offsetof(IterParams, neg) == 4, "the generated walks read pos, neg");
static_assert(sizeof(WalkArgs) == 11 * sizeof(void*) + 24, "WalkArgs layout");

// Declarations every generated kernel uses. The operations follow
// up_reg_walk in engine.hip: pos0 and recip are its helpers
// (crates/solver/src/cfr.rs:113-121), readlane its record reads; enc16 is
// its regret encoding (storage.rs:274-281: y = v * (MAX / max), rounded half
// away from zero, Rust's saturating `as i16`) and wmax its wave maximum.
static const char* kPreamble = R"(#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
typedef unsigned int u32;
typedef unsigned long long u64;
struct __attribute__((aligned(16))) H4 { u32 x, y, z, w; };
struct WalkArgs {
  const u32* list;
  const H4* seg_head;
  const unsigned char* mrec;
  const unsigned char* malive;
  float* vals;
  float* regrets;
  const float* prm;
  const float* zrow;
  short* rq;
  const float* rsc;
  int* rnx;
  u32 i0, nseg, tiles, rowlen, v0, vb;
};
__device__ __forceinline__ float pos0(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float recip(float d) {
  const float i = 1.0f / d;
  return (d > 0.0f && __builtin_isfinite(i)) ? i : 0.0f;
}
__device__ __forceinline__ u32 rl32(u32 x, u32 m) { return (u32)__builtin_amdgcn_readlane((int)x, (int)m); }
__device__ __forceinline__ u64 rl64(u64 x, u32 m) { return (u64)rl32((u32)x, m) | (u64)rl32((u32)(x >> 32), m) << 32; }
__device__ __forceinline__ short enc16(float v, float inv) {
  const float y = v * inv;
  const float z = y + __builtin_copysignf(0.5f, y);
  if (!(z == z)) return 0;
  return (short)(int)__builtin_fminf(__builtin_fmaxf(z, -32768.0f), 32767.0f);
}
__device__ __forceinline__ float dmax(float x, int src) { return __builtin_fmaxf(x, __builtin_bit_cast(float, src)); }
__device__ __forceinline__ float wmax(float x) {
  x = dmax(x, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x111, 0xF, 0xF, false));
  x = dmax(x, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x112, 0xF, 0xF, false));
  x = dmax(x, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x114, 0xF, 0xF, false));
  x = dmax(x, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x118, 0xF, 0xF, false));
  x = dmax(x, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x142, 0xA, 0xF, false));
  x = dmax(x, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x143, 0xC, 0xF, false));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
}
)";

static std::string kernel_name(size_t i, int p, int kind) {
  static const char* pre[kWalkKinds] = {"pts_walk_", "pts_best_", "pts_wmax_", "pts_wenc_"};
  return std::string(pre[kind]) + std::to_string(i) + "_p" + std::to_string(p);
}

// One kernel: the walk of shape S for traverser p, of the given kind: the
// f32 DCFR walk (regret update), the FP64 best response (no storage), or one
// of the two 16-bit DCFR passes (the largest new regret of each own node,
// then the new regrets encoded with the new scale). Member m's value is v<m>;
// own node m's regret row for action a is r<m>_<a>. Lane m (and lane m - 64
// of the second set) holds member m's record; LIVE, FROZEN, VS, OF and PS
// read member m's flags, value row, storage offset and scale slot.
static void emit_kernel(std::ostringstream& s, const Shape& S, size_t idx, int p, int kind) {
  const uint32_t M = (uint32_t)S.size();
  const bool two = M > 64;
  const bool best = kind == kWalkBest, q16 = kind == kWalkMax16 || kind == kWalkEnc16;
  const char* T = best ? "double" : "float";
  const char* zero = best ? "0.0" : "0.0f";
  auto live = [&](uint32_t m) {
    return m < 64 ? "((L0 >> " + std::to_string(m) + ") & 1ull)" : "((L1 >> " + std::to_string(m - 64) + ") & 1ull)";
  };
  auto frozen = [&](uint32_t m) {
    return m < 64 ? "((F0 >> " + std::to_string(m) + ") & 1ull)" : "((F1 >> " + std::to_string(m - 64) + ") & 1ull)";
  };
  auto vs = [&](uint32_t m) {
    return m < 64 ? "rl32(vs0, " + std::to_string(m) + "u)" : "rl32(vs1, " + std::to_string(m - 64) + "u)";
  };
  auto of = [&](uint32_t m) {
    return m < 64 ? "rl64(of0, " + std::to_string(m) + "u)" : "rl64(of1, " + std::to_string(m - 64) + "u)";
  };
  auto ps = [&](uint32_t m) {
    return m < 64 ? "rl32(ps0, " + std::to_string(m) + "u)" : "rl32(ps1, " + std::to_string(m - 64) + "u)";
  };
  auto v = [&](uint32_t m) { return "v" + std::to_string(m); };
  s << "extern \"C\" __global__ void __launch_bounds__(64, 4) " << kernel_name(idx, p, kind) << "(WalkArgs A) {\n";
  // workgroup L runs on XCD L % 8; its slot L / 8 there names (segment,
  // hand tile) with a segment's tiles consecutive, so one XCD walks all of a
  // segment's tiles and its member records come from that XCD's L2
  s << "  const u32 xcd = blockIdx.x & 7u, slot = blockIdx.x >> 3;\n";
  s << "  const u32 sl = slot / A.tiles, tile = slot - sl * A.tiles;\n";
  s << "  const u32 si = sl * 8u + xcd;\n";
  s << "  if (si >= A.nseg) return;\n";
  s << "  const u32 seg = A.list[A.i0 + si];\n";
  s << "  const H4 head = A.seg_head[seg];\n";
  s << "  const u32 moff = head.x;\n";
  s << "  const u32 np = " << (p ? "head.w >> 16" : "head.w & 0xFFFFu") << ";\n";
  s << "  const u32 lane = threadIdx.x;\n";
  s << "  const u32 h = tile * 64u + lane;\n";
  s << "  const bool act = h < np;\n";
  s << "  const u32 hh = act ? h : 0u;\n";
  s << "  " << T << "* const vals = (" << T << "*)A.vals;\n";
  s << "  const " << T << "* const zrow = (const " << T << "*)A.zrow;\n";
  if (!best) s << "  const float pos = A.prm[0], neg = A.prm[1];\n";
  for (int q = 0; q < (two ? 2 : 1); q++) {
    const uint32_t cnt = q == 0 ? (M < 64 ? M : 64) : M - 64;
    s << "  u32 vs" << q << " = 0u, fz" << q << " = 0u, lv" << q << " = 0u, ps" << q << " = 0u;\n";
    s << "  u64 of" << q << " = 0ull;\n";
    s << "  if (lane < " << cnt << "u) {\n";
    s << "    const u32 mm = moff + " << 64 * q << "u + lane;\n";
    s << "    const unsigned char* r = A.mrec + (u64)mm * 48u;\n";
    s << "    of" << q << " = *(const u64*)r;\n";
    s << "    vs" << q << " = *(const u32*)(r + 12);\n";
    if (q16) s << "    ps" << q << " = *(const u32*)(r + 44);\n";
    s << "    fz" << q << " = r[41];\n";
    s << "    lv" << q << " = A.malive[mm];\n";
    s << "  }\n";
    s << "  vs" << q << " = vs" << q << " < A.v0 ? vs" << q << " : vs" << q << " - A.vb;  // the group's value row (Bias::v)\n";
    s << "  const u64 L" << q << " = __builtin_amdgcn_ballot_w64(lv" << q << " != 0u);\n";
    if (!best) s << "  const u64 F" << q << " = __builtin_amdgcn_ballot_w64((fz" << q << " & 8u) != 0u);\n";
  }
  // every row the walk reads, all loads issued before the walk and none
  // behind a branch: the values of chance / terminal / leaf members and, in
  // the DCFR walks, the regret rows of own nodes (up_reg_walk's preload,
  // 16-bit regrets decoded as i16 * (scale / 32767)); a dead member reads
  // the zero row instead, so it holds 0 as in the generic walk
  for (uint32_t m = 0; m < M; m++) {
    const Member& x = S[m];
    if (x.kind >= 2) {
      s << "  const " << T << " " << v(m) << "l = (" << live(m) << " ? vals + (u64)" << vs(m) << " * A.rowlen : zrow)[hh];\n";
    } else if ((int)x.kind == p && !best && !q16) {
      s << "  const float* q" << m << " = " << live(m) << " ? A.regrets + " << of(m) << " : A.zrow;\n";
      for (uint32_t a = 0; a < x.na; a++)
        s << "  const float r" << m << "_" << a << " = q" << m << "[(u64)" << a << "u * np + hh];\n";
    } else if ((int)x.kind == p && q16) {
      s << "  const short* q" << m << " = " << live(m) << " ? A.rq + " << of(m) << " : (const short*)A.zrow;\n";
      s << "  const float st" << m << " = " << live(m) << " ? A.rsc[" << ps(m) << "] / 32767.0f : 0.0f;\n";
      for (uint32_t a = 0; a < x.na; a++)
        s << "  const float r" << m << "_" << a << " = (float)q" << m << "[(u64)" << a << "u * np + hh] * st" << m << ";\n";
    }
  }
  for (uint32_t m = 0; m < M; m++)
    if (S[m].kind >= 2) s << "  const " << T << " " << v(m) << " = " << v(m) << "l;\n";
  // the walk, members in reverse breadth-first order; a dead decision node
  // is 0 and updates nothing
  std::ostringstream tail;
  for (uint32_t m = M; m-- > 0;) {
    const Member& x = S[m];
    if (x.kind >= 2) continue;
    if (!best && (int)x.kind == p)
      for (uint32_t a = 0; a < x.na; a++) s << "  float u" << m << "_" << a << ";\n";
    s << "  " << T << " " << v(m) << ";\n  {\n";
    if ((int)x.kind != p) {  // opponent node: the sum of its children
      s << "    " << T << " t = " << zero << ";\n";
      for (uint32_t a = 0; a < x.na; a++) s << "    t += " << v(x.cl + a) << ";\n";
      s << "    " << v(m) << " = " << live(m) << " ? t : " << zero << ";\n";
    } else if (best) {  // own node, best response: the larger child value, the first on ties
      s << "    double b = " << v(x.cl) << ";\n";
      for (uint32_t a = 1; a < x.na; a++) {
        const std::string c = v(x.cl + a);
        s << "    b = b > " << c << " ? b : (" << c << " > b ? " << c << " : b);\n";
      }
      s << "    " << v(m) << " = " << live(m) << " ? b : 0.0;\n";
    } else {  // own node: regret matching, then the regret update with the DCFR discount (cfr.rs:315-382)
      s << "    float d = 0.0f, o = 0.0f;\n";
      for (uint32_t a = 0; a < x.na; a++)
        s << "    { const float r = pos0(r" << m << "_" << a << "); d += r; o += r * " << v(x.cl + a) << "; }\n";
      s << "    const float inv = recip(d);\n";
      s << "    o *= inv;\n";
      s << "    float sm = 0.0f;\n";
      for (uint32_t a = 0; a < x.na; a++) s << "    sm += " << v(x.cl + a) << ";\n";
      s << "    const float uo = sm * (1.0f / (float)" << (uint32_t)x.na << "u);\n";
      s << "    o = inv == 0.0f ? uo : o;\n";
      for (uint32_t a = 0; a < x.na; a++) {
        const std::string r = "r" + std::to_string(m) + "_" + std::to_string(a);
        s << "    u" << m << "_" << a << " = " << r << " * (" << r << " > 0.0f ? pos : neg) + " << v(x.cl + a) << " - o;\n";
      }
      s << "    " << v(m) << " = " << live(m) << " ? o : 0.0f;\n";
      // the updated regrets are stored after the walk: a store followed by
      // a read of a loaded row would make that read wait for the store
      if (kind == kWalkF32) {
        tail << "  if (" << live(m) << " && !" << frozen(m) << " && act) {\n";
        tail << "    float* w = A.regrets + " << of(m) << ";\n";
        for (uint32_t a = 0; a < x.na; a++) tail << "    w[(u64)" << a << "u * np + h] = u" << m << "_" << a << ";\n";
        tail << "  }\n";
      } else if (kind == kWalkMax16) {  // the node's largest |new regret| over its hands, as float bits
        tail << "  if (" << live(m) << " && !" << frozen(m) << ") {\n";
        tail << "    float lm = 0.0f;\n";
        for (uint32_t a = 0; a < x.na; a++) tail << "    if (act) lm = __builtin_fmaxf(lm, __builtin_fabsf(u" << m << "_" << a << "));\n";
        tail << "    lm = wmax(lm);\n";
        tail << "    if (lane == 0u) atomicMax(A.rnx + " << ps(m) << ", __builtin_bit_cast(int, lm));\n";
        tail << "  }\n";
      } else {  // kWalkEnc16: the new regrets encoded with the node's new scale
        tail << "  if (" << live(m) << " && !" << frozen(m) << ") {\n";
        tail << "    const float mx = __builtin_bit_cast(float, A.rnx[" << ps(m) << "]);\n";
        tail << "    const float inv = mx > 0.0f ? 32767.0f / mx : 0.0f;\n";
        tail << "    if (act) {\n";
        tail << "      short* w = A.rq + " << of(m) << ";\n";
        for (uint32_t a = 0; a < x.na; a++)
          tail << "      w[(u64)" << a << "u * np + h] = mx > 0.0f ? enc16(u" << m << "_" << a << ", inv) : (short)0;\n";
        tail << "    }\n";
        tail << "  }\n";
      }
    }
    s << "  }\n";
  }
  s << tail.str();
  // the root's value row (the 16-bit encoding pass leaves it to the first pass)
  if (kind != kWalkEnc16) s << "  if (act) vals[(u64)" << vs(0) << " * A.rowlen + h] = v0;\n";
  s << "}\n";
}

// the kinds an engine launches: the f32 walk and the best response, or with
// 16-bit storage the best response and the two 16-bit passes
static bool kind_used(int kind, bool q16) { return q16 ? kind != kWalkF32 : kind == kWalkF32 || kind == kWalkBest; }

std::string walk_source(const std::vector<Shape>& shapes, bool q16) {
  std::ostringstream s;
  s << kPreamble;
  for (size_t i = 0; i < shapes.size(); i++)
    for (int kind = 0; kind < kWalkKinds; kind++)
      if (kind_used(kind, q16))
        for (int p = 0; p < 2; p++) emit_kernel(s, shapes[i], i, p, kind);
  return s.str();
}

namespace {
std::mutex g_mu;
std::map<std::string, hipModule_t> g_modules;  // (device, source) -> module, kept for the process
}  // namespace

#define JITCHK(x)                                                                               \
  do {                                                                                          \
    hipError_t e_ = (x);                                                                        \
    if (e_ != hipSuccess) {                                                                     \
      set_error(std::string("HIP error in ") + #x + ": " + hipGetErrorString(e_));              \
      set_error_detail(ErrorDetail{(int32_t)e_, 0, 0});                                         \
      return e_ == hipErrorOutOfMemory ? kOutOfMemory : kDriver;                                \
    }                                                                                           \
  } while (0)

Status compile_walks(const std::vector<Shape>& shapes, bool q16, std::vector<hipFunction_t>& fn, double* seconds) {
  if (seconds) *seconds = 0.0;
  fn.clear();
  if (shapes.empty()) return kOk;
  for (const Shape& S : shapes) {
    if (S.empty() || S.size() > kMaxMembers) {
      set_error("invalid input: a specialised walk takes 1 to 80 members");
      return kInvalidInput;
    }
    for (size_t m = 0; m < S.size(); m++)
      if (S[m].kind < 2 && (S[m].na == 0 || S[m].cl <= m || (size_t)S[m].cl + S[m].na > S.size())) {
        set_error("invalid input: a decision member's children must follow it in the segment");
        return kInvalidInput;
      }
  }
  int dev = 0;
  JITCHK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  JITCHK(hipGetDeviceProperties(&prop, dev));
  std::string arch = prop.gcnArchName;
  const size_t colon = arch.find(':');
  if (colon != std::string::npos) arch = arch.substr(0, colon);
  const std::string src = walk_source(shapes, q16);
  const std::string key = std::to_string(dev) + "\n" + arch + "\n" + src;
  std::lock_guard<std::mutex> lk(g_mu);
  hipModule_t mod = nullptr;
  auto it = g_modules.find(key);
  if (it != g_modules.end()) {
    mod = it->second;
  } else {
    const auto t0 = std::chrono::steady_clock::now();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "pts_walks.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
      set_error("hiprtc: could not create the walk program");
      return kDriver;
    }
    // the library's own device flags: no contraction, IEEE division and
    // denormals, packed FP32 off
    const std::string off = "--offload-arch=" + arch;
    const char* opts[] = {off.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (rc != HIPRTC_SUCCESS) {
      size_t n = 0;
      (void)hiprtcGetProgramLogSize(prog, &n);
      std::string log(n, '\0');
      if (n) (void)hiprtcGetProgramLog(prog, &log[0]);
      (void)hiprtcDestroyProgram(&prog);
      set_error("hiprtc: the walk program did not compile: " + log.substr(0, 2000));
      return kDriver;
    }
    size_t n = 0;
    (void)hiprtcGetCodeSize(prog, &n);
    std::vector<char> code(n);
    (void)hiprtcGetCode(prog, code.data());
    (void)hiprtcDestroyProgram(&prog);
    JITCHK(hipModuleLoadData(&mod, code.data()));
    g_modules[key] = mod;
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  fn.assign(shapes.size() * kWalkKinds * 2, nullptr);
  for (size_t i = 0; i < shapes.size(); i++)
    for (int kind = 0; kind < kWalkKinds; kind++)
      if (kind_used(kind, q16))
        for (int p = 0; p < 2; p++)
          JITCHK(hipModuleGetFunction(&fn[walk_fn(i, kind, p)], mod, kernel_name(i, p, kind).c_str()));
  return kOk;
}

}  // namespace jit
}  // namespace pts
