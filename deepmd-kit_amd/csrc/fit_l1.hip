// Fitting layer 1 fused with the descriptor, forward and backward, for the
// split-FP16 precision: neither the descriptor D nor dE/dD is written to
// memory.
//
// Reference math (deepmd/pt/model/descriptor/se_a.py:844-850,
// deepmd/pt/model/task/fitting.py:734-899):
//   D[q]  = sum_c xyz[c][k] xyz[c][m],   q = 16 k + m  (k < d3, m < 16)
//   z1[h] = sum_q D[q] W1[q][h]                        (h < 256; the tail adds b1)
//   GD[q] = sum_h gz1[h] W1[q][h]                      (dE/dD)
//   gxyz[c][p] = sum_m GD[16 p + m] xyz[c][m] + [p < 16] sum_k GD[16 k + p] xyz[c][k]
//
// MI355X mapping (v_mfma_f32_16x16x32_f16, operands split into hi + lo FP16
// pairs against FP16 weights, FP32 accumulation).  A workgroup of 8
// wavefronts takes 64 slots: four groups of 16 slots, two wavefronts per
// group.  Lane (g = lane / 16, n = lane % 16) works for slot n of its group.
//  * Pipeline (both kernels): stage s covers the channels k = 2 s, 2 s + 1 --
//    16 KB of W1 fragments plus the workgroup's xyz[slot][c][k] -- and is
//    copied straight from global memory into LDS (global_load_lds, no
//    registers), through a ring of three stage buffers: stage s + 2 is issued
//    while s is consumed, with one s_waitcnt + s_barrier per stage.
//  * Forward: Z^T = W1^T D^T with the 256 outputs h as M (each wavefront of a
//    pair owns 128), the slots as N and q as K in chunks of 32 (one stage).
//    The B operand -- 8 values D[slot][32 s + 8 g + j] of the lane -- is
//    formed in registers from xyz (4 FMAs a value) and split.  z1 leaves as
//    FP32 rows of 256 for the fitting tail.
//  * Backward: GD^T = W1 gz1^T with q as M (one 16-row block per channel k),
//    the slots as N and h as K; the pair splits h, each wavefront keeping its
//    half of gz1 (hi and lo) in registers for the whole kernel, and the two
//    partial sums of a block meet in LDS.  Each block's output,
//    GD[slot][k][m = 4 g + i], is contracted at once: the first term over m
//    (lanes g and registers i, reduced by lane permutes) gives dE/dxyz[c][k];
//    the second term accumulates sum_k GD[k][m] xyz[c][k] per lane, and the
//    pair's partial sums meet in LDS at the end and are added to the
//    channels p < 16.
#include "dpk_common.h"
#include "half16.h"

namespace dpk {
namespace {
namespace l1 {

using h16::v2u;
using h16::v4f;
using h16::v4u;

constexpr int kH = 256;                       // layer-1 width (padded)
constexpr int kHB = kH / 16;                  // output blocks of the forward
constexpr int kHC = kH / 32;                  // K chunks of the backward
constexpr int kWaves = 8;
constexpr int kThreads = 64 * kWaves;
constexpr int kSlots = 16 * (kWaves / 2);     // slots per workgroup
constexpr int kFrag = 64 * 16;                // one A fragment (64 lanes x 8 halves)
constexpr int kStage = kHB * kFrag;           // 16 KB of W1 fragments per pipeline stage
constexpr int kStageI4 = kStage / 16;         // the same in 16-byte pieces (2 per thread)
constexpr int kBwdBlock = kHC * kFrag;        // 8 KB: all K chunks of one q block

// Pipeline stages: stage s covers the descriptor channels k = 2 s, 2 s + 1 --
// forward K chunk s (q = 32 s .. 32 s + 31) and backward q blocks 2 s, 2 s + 1.
__host__ __device__ inline int n_stages(int d3) { return (d3 + 1) / 2; }
__host__ __device__ inline size_t fwd_bytes(int d3) { return (size_t)n_stages(d3) * kStage; }
__host__ __device__ inline size_t bwd_bytes(int d3) { return (size_t)n_stages(d3) * kStage; }

typedef __attribute__((address_space(1))) void gmem_t;
typedef __attribute__((address_space(3))) void lmem_t;

// s_waitcnt immediates (vmcnt in bits 3:0, expcnt 6:4, lgkmcnt 11:8)
constexpr int kWaitVm3Lgkm0 = 0x0073;  // vmcnt(3) expcnt(7) lgkmcnt(0)
constexpr int kWaitVm0Lgkm0 = 0x0070;  // vmcnt(0) expcnt(7) lgkmcnt(0)

template <class E>
__device__ __forceinline__ v4u pack8(const float* a) {
  return v4u{E::pack2(a[0], a[1]), E::pack2(a[2], a[3]), E::pack2(a[4], a[5]),
             E::pack2(a[6], a[7])};
}

// lo = fp16(x - hi) of eight values packed in hi
__device__ __forceinline__ v4u lo8(const float* a, v4u hi) {
  v4u r;
#pragma unroll
  for (int w = 0; w < 4; ++w)
    r[w] = h16::F16::pack2(h16::sub_f16_lo(a[2 * w], hi[w]), h16::sub_f16_hi(a[2 * w + 1], hi[w]));
  return r;
}

// *p where ok, else 0: the load is unconditional (p must be a valid address
// either way) so that no branch splits the wavefront's load stream and the
// wait counts of software-pipelined loads stay exact
__device__ __forceinline__ float ld_or0(const float* p, bool ok) {
  const float v = *p;
  return ok ? v : 0.0f;
}

// Row sums of a wavefront (rows of 16 lanes, g = lane / 16) scattered over the
// rows: returns sum over the four rows of p[g] at every lane of row g.  Two
// v_permlane32_swap (rows g, g ^ 2) and one v_permlane16_swap (rows g, g ^ 1):
// swap(a, b) exchanges the upper rows of a with the lower rows of b, so that
// the sum of the two results holds, at each lane, its own value of one operand
// plus the partner's value of the same operand.
__device__ __forceinline__ float row_sums_scatter(const float p[4]) {
  const auto s02 = __builtin_amdgcn_permlane32_swap(__float_as_uint(p[0]), __float_as_uint(p[2]),
                                                    false, false);
  const auto s13 = __builtin_amdgcn_permlane32_swap(__float_as_uint(p[1]), __float_as_uint(p[3]),
                                                    false, false);
  // rows 0, 1: p[0] and p[1] over rows {g, g ^ 2}; rows 2, 3: p[2] and p[3]
  const float q0 = __uint_as_float(s02[0]) + __uint_as_float(s02[1]);
  const float q1 = __uint_as_float(s13[0]) + __uint_as_float(s13[1]);
  const auto t = __builtin_amdgcn_permlane16_swap(__float_as_uint(q0), __float_as_uint(q1), false,
                                                  false);
  return __uint_as_float(t[0]) + __uint_as_float(t[1]);
}

// Image (FP16): forward part, entry ((kc * 16 + mb) * 64 + lane) * 8 + j =
// W1[32 kc + 8 g + j][16 mb + m]; backward part, entry ((b * 8 + kc) * 64 +
// lane) * 8 + j = W1[16 b + m][32 kc + 8 g + j] (lane = 16 g + m; rows q >= 16
// d3 zero).
__global__ void k_pack(unsigned short* __restrict__ img, const float* __restrict__ w1, int d3) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nf = (long)fwd_bytes(d3) / 2, nb = (long)bwd_bytes(d3) / 2;
  if (e >= nf + nb) return;
  const int din = 16 * d3;
  int q, h;
  if (e < nf) {
    const long f = e / 8;
    const int j = (int)(e % 8), lane = (int)(f % 64), mb = (int)((f / 64) % kHB), kc = (int)(f / 64 / kHB);
    q = 32 * kc + 8 * (lane >> 4) + j;
    h = 16 * mb + (lane & 15);
  } else {
    const long e2 = e - nf, f = e2 / 8;
    const int j = (int)(e2 % 8), lane = (int)(f % 64), kc = (int)((f / 64) % kHC), b = (int)(f / 64 / kHC);
    q = 16 * b + (lane & 15);
    h = 32 * kc + 8 * (lane >> 4) + j;
  }
  const float v = q < din ? w1[(size_t)q * kH + h] : 0.0f;
  img[e] = h16::F16::from_f(v);
}

// D[slot][32 s + 8 g + j] for the lane's eight j from xk = xyz[c][2 s + g / 2]
// and xm = xyz[c][8 (g & 1) + j]
__device__ __forceinline__ void d_chunk(v4f xk, const float xm[4][8], float dv[8]) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    float a = xk[0] * xm[0][j];
    a = __builtin_fmaf(xk[1], xm[1][j], a);
    a = __builtin_fmaf(xk[2], xm[2][j], a);
    dv[j] = __builtin_fmaf(xk[3], xm[3][j], a);
  }
}

// LDS byte address of a workgroup-shared pointer
__device__ __forceinline__ unsigned lds_addr(const void* p) {
  return (unsigned)(size_t)(const __attribute__((address_space(3))) void*)p;
}

// Eight 16-byte LDS reads at a + O0..O7 and one at x, completed before
// returning.  Written out so that the compiler's wait-count pass, which
// cannot tell the stage buffers' direct-to-LDS copies apart across the
// pipelined loop and would wait for every copy in flight before each read,
// does not see them; the kernels wait for exactly the stage they read
// (s_waitcnt vmcnt + s_barrier) before calling this.
template <int O0, int O1, int O2, int O3, int O4, int O5, int O6, int O7>
__device__ __forceinline__ void lds_read9(v4u (&f)[8], v4f& x, unsigned a, unsigned xa) {
  asm volatile(
      "ds_read_b128 %0, %9 offset:%c10\n\t"
      "ds_read_b128 %1, %9 offset:%c11\n\t"
      "ds_read_b128 %2, %9 offset:%c12\n\t"
      "ds_read_b128 %3, %9 offset:%c13\n\t"
      "ds_read_b128 %4, %9 offset:%c14\n\t"
      "ds_read_b128 %5, %9 offset:%c15\n\t"
      "ds_read_b128 %6, %9 offset:%c16\n\t"
      "ds_read_b128 %7, %9 offset:%c17\n\t"
      "ds_read_b128 %8, %18\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(f[0]), "=&v"(f[1]), "=&v"(f[2]), "=&v"(f[3]), "=&v"(f[4]), "=&v"(f[5]),
        "=&v"(f[6]), "=&v"(f[7]), "=&v"(x)
      : "v"(a), "i"(O0), "i"(O1), "i"(O2), "i"(O3), "i"(O4), "i"(O5), "i"(O6), "i"(O7), "v"(xa)
      : "memory");
}

// Four 16-byte LDS reads at a + O0..O3, completed before returning (as above)
template <int O0, int O1, int O2, int O3>
__device__ __forceinline__ void lds_read4(v4u (&f)[4], unsigned a) {
  asm volatile(
      "ds_read_b128 %0, %4 offset:%c5\n\t"
      "ds_read_b128 %1, %4 offset:%c6\n\t"
      "ds_read_b128 %2, %4 offset:%c7\n\t"
      "ds_read_b128 %3, %4 offset:%c8\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(f[0]), "=&v"(f[1]), "=&v"(f[2]), "=&v"(f[3])
      : "v"(a), "i"(O0), "i"(O1), "i"(O2), "i"(O3)
      : "memory");
}

__device__ __forceinline__ v4f lds_read1(unsigned a) {
  v4f x;
  asm volatile("ds_read_b128 %0, %1\n\ts_waitcnt lgkmcnt(0)" : "=v"(x) : "v"(a) : "memory");
  return x;
}

// Per-thread source of the pipeline's direct-to-LDS copies.  Each stage holds
// 16 KB of W1 fragments (two 16-byte pieces per thread) and the 512 values
// xyz[slot][c][2 s + kk] of the workgroup's 64 slots (one per thread: thread
// t = 8 slot + 4 kk + c), laid out so that a lane reads its slot's four c of
// one kk as one 16-byte LDS read.
struct StageSrc {
  const int4* w1;    // image piece of this thread in stage 0
  const float* row;  // xyz[atom of the thread's slot][c]
  int kk, nstage, d3;

  __device__ __forceinline__ StageSrc(const unsigned char* img, const float* xyz, const int* order,
                                      int n, int d3_) {
    const int t = threadIdx.x;
    const int slot = blockIdx.x * kSlots + (t >> 3);
    const int atom = slot < n ? order[slot] : 0;  // any existing row for padding slots
    w1 = reinterpret_cast<const int4*>(img) + t;
    row = xyz + (size_t)atom * 4 * d3_ + (size_t)(t & 3) * d3_;
    kk = (t >> 2) & 1;
    nstage = n_stages(d3_);
    d3 = d3_;
  }
  // three vector-memory operations per thread, always issued (stages past
  // the end reload the last one) so that the wait counts stay exact
  __device__ __forceinline__ void issue(int s, int4* w1_lds, float* xk_lds) const {
    const int st = s < nstage ? s : nstage - 1;
    const int4* src = w1 + (size_t)st * kStageI4;
    const int wave = threadIdx.x >> 6;
    __builtin_amdgcn_global_load_lds((gmem_t*)src, (lmem_t*)(w1_lds + 64 * wave), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((gmem_t*)(src + kThreads),
                                     (lmem_t*)(w1_lds + kThreads + 64 * wave), 16, 0, 0);
    const int k = 2 * st + kk;
    __builtin_amdgcn_global_load_lds((gmem_t*)(row + (k < d3 ? k : d3 - 1)),
                                     (lmem_t*)(xk_lds + 64 * wave), 4, 0, 0);
  }
};

// Forward.  Three-stage ring of direct-to-LDS copies: at step s the copies of
// stage s + 1 are in flight and those of stage s + 2 are issued, so each
// copy has two steps to land.  Every stage buffer is its own LDS object and
// the loop is unrolled by three, which lets the compiler see that reads of
// one stage never alias the copies into another.
template <class E>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) k_fwd(float* __restrict__ z1,
                                                     const float* __restrict__ xyz,
                                                     const int* __restrict__ order,
                                                     const unsigned char* __restrict__ img, int n,
                                                     int d3) {
  __shared__ int4 w1s0[kStageI4], w1s1[kStageI4], w1s2[kStageI4];
  __shared__ float xks0[kThreads], xks1[kThreads], xks2[kThreads];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int half = wave & 1;
  const int sl = 16 * (wave >> 1) + c;  // slot within the workgroup
  const int slot = blockIdx.x * kSlots + sl;
  const bool valid = slot < n;
  const float* xa = xyz + (size_t)(valid ? order[slot] : 0) * 4 * d3;
  // xyz[c][m] for the lane's eight m = 8 (g & 1) + j (m < 16 <= d3)
  float xm[4][8];
#pragma unroll
  for (int cc = 0; cc < 4; ++cc)
#pragma unroll
    for (int j = 0; j < 8; ++j) xm[cc][j] = ld_or0(xa + cc * d3 + 8 * (g & 1) + j, valid);
  v4f acc[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) acc[b] = v4f{0.f, 0.f, 0.f, 0.f};
  const StageSrc src(img, xyz, order, n, d3);
  const int nstage = src.nstage;
  src.issue(0, w1s0, xks0);
  src.issue(1, w1s1, xks1);
#define DPK_L1_FWD_STEP(S, W1C, XKC, W1N, XKN)                                          \
  if ((S) < nstage) {                                                                  \
    /* this wave's copies of stage S have landed (those of S + 1 may not) ... */       \
    __builtin_amdgcn_s_waitcnt(kWaitVm3Lgkm0);                                         \
    /* ... and every wave's; the stage read at step S - 1 is free again */             \
    __builtin_amdgcn_s_barrier();                                                      \
    src.issue((S) + 2, W1N, XKN);                                                      \
    v4u af[8];                                                                         \
    v4f xk;                                                                            \
    lds_read9<0, 1024, 2048, 3072, 4096, 5120, 6144, 7168>(                            \
        af, xk, lds_addr(W1C + (8 * half) * 64 + lane),                                \
        lds_addr(XKC + 8 * sl + 4 * (g >> 1)));                                        \
    float dv[8];                                                                       \
    d_chunk(xk, xm, dv);                                                               \
    const v4u bh = pack8<E>(dv);                                                       \
    v4u bl;                                                                            \
    if constexpr (E::kSplit) bl = lo8(dv, bh);                                         \
    _Pragma("unroll") for (int b = 0; b < 8; ++b) {                                    \
      acc[b] = E::mfma32(af[b], bh, acc[b]);                                           \
      if constexpr (E::kSplit) acc[b] = E::mfma32(af[b], bl, acc[b]);                  \
    }                                                                                  \
  }
  for (int s = 0; s < nstage; s += 3) {
    DPK_L1_FWD_STEP(s, w1s0, xks0, w1s2, xks2)
    DPK_L1_FWD_STEP(s + 1, w1s1, xks1, w1s0, xks0)
    DPK_L1_FWD_STEP(s + 2, w1s2, xks2, w1s1, xks1)
  }
#undef DPK_L1_FWD_STEP
  // no copy may still be in flight when the workgroup's LDS is handed on
  __builtin_amdgcn_s_waitcnt(kWaitVm0Lgkm0);
  if (valid) {
    float* o = z1 + (size_t)slot * kH + 4 * g;
#pragma unroll
    for (int b = 0; b < 8; ++b)
      *reinterpret_cast<v4f*>(o + 16 * (8 * half + b)) = acc[b];
  }
}

// Backward.  The pair splits h: wavefront `half` holds gz1 for h in [128 half,
// 128 half + 128) (hi and lo: 32 registers) and forms partial GD over its h
// for both q blocks of the stage; it keeps the partial of block 2 s + half and
// hands the other to its partner through LDS (a ring of two exchange
// buffers).  The exchange is read one step later, after the next barrier,
// when the block is finished: first term and the second term's accumulation.
// The first-term outputs gather in LDS (a ring of three 2 KB tiles, one per
// step) and leave two steps at a time as 16-byte runs of four channels per
// (slot, component): one store per 4 channels instead of one per channel.
template <class E>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) k_bwd(
    float* __restrict__ gxyz, const unsigned short* __restrict__ gz1,
    const float* __restrict__ xyz, const int* __restrict__ order,
    const unsigned char* __restrict__ img, int n, int d3) {
  __shared__ int4 w1s0[kStageI4], w1s1[kStageI4], w1s2[kStageI4];
  __shared__ float xks0[kThreads], xks1[kThreads], xks2[kThreads];
  __shared__ v4f xch0[kThreads], xch1[kThreads];          // partner partials
  __shared__ float2 til0[256], til1[256], til2[256];       // [slot][c] (k = 2 s, 2 s + 1)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int half = wave & 1, grp = wave >> 1;
  const int sl = 16 * grp + c;
  const int slot = blockIdx.x * kSlots + sl;
  const bool valid = slot < n;
  const float* xa = xyz + (size_t)(valid ? order[slot] : 0) * 4 * d3;
  // B operand: the slot's gz1 over this wavefront's h half, hi and lo parts
  constexpr int kStride = E::kSplit ? 2 * kH : kH;
  constexpr int kHalfC = kHC / 2;  // K chunks of 32 h per wavefront
  v4u gh[kHalfC], gl[kHalfC];
  const unsigned short* gr = gz1 + (size_t)(valid ? slot : 0) * kStride + 128 * half + 8 * g;
#pragma unroll
  for (int kc = 0; kc < kHalfC; ++kc) {
    const v4u h = *reinterpret_cast<const v4u*>(gr + 32 * kc);
    gh[kc] = valid ? h : v4u{0u, 0u, 0u, 0u};
    if constexpr (E::kSplit) {
      const v4u l = *reinterpret_cast<const v4u*>(gr + kH + 32 * kc);
      gl[kc] = valid ? l : v4u{0u, 0u, 0u, 0u};
    }
  }
  // xyz[c][m] for the lane's m = 4 g + i (first term)
  float xm[4][4];
#pragma unroll
  for (int cc = 0; cc < 4; ++cc)
#pragma unroll
    for (int i = 0; i < 4; ++i) xm[cc][i] = ld_or0(xa + cc * d3 + 4 * g + i, valid);
  float acc2[4][4];  // sum_k GD[k][m = 4 g + i] xyz[c][k] over the wavefront's blocks
#pragma unroll
  for (int cc = 0; cc < 4; ++cc)
#pragma unroll
    for (int i = 0; i < 4; ++i) acc2[cc][i] = 0.0f;
  const StageSrc src(img, xyz, order, n, d3);
  const int nstage = src.nstage;
  // Output runs: thread t with (t / 4) even writes (slot t / 8, component
  // t % 4) -- the same row as its stage copies of xyz.
  const bool writer = ((threadIdx.x >> 2) & 1) == 0 &&
                      blockIdx.x * kSlots + (int)(threadIdx.x >> 3) < n;
  float* const grow = gxyz + (src.row - xyz);
  const int tix = ((threadIdx.x >> 3) * 4 + (threadIdx.x & 3));  // tile index of the run
  // channels k0 .. k0 + 3 from the tiles of steps k0 / 2 and k0 / 2 + 1
  auto put4 = [&](int k0, const float2* t0, const float2* t1) __attribute__((always_inline)) {
    const float2 a = t0[tix], b = t1[tix];
    if (writer) {
      if ((d3 & 3) == 0) {
        *reinterpret_cast<v4f*>(grow + k0) = v4f{a.x, a.y, b.x, b.y};
      } else {
        if (k0 < d3) grow[k0] = a.x;
        if (k0 + 1 < d3) grow[k0 + 1] = a.y;
        if (k0 + 2 < d3) grow[k0 + 2] = b.x;
        if (k0 + 3 < d3) grow[k0 + 3] = b.y;
      }
    }
  };
  // the block finished one step late: own partial GD and xyz[c][k]
  v4f own = {0.f, 0.f, 0.f, 0.f}, xkp = {0.f, 0.f, 0.f, 0.f};
  // lane (g, slot) of wavefront half finishes channel k = 2 s + half
  auto finish = [&](const v4f* xr, float2* tw) __attribute__((always_inline)) {
    const v4f gd = own + xr[64 * wave + lane];  // GD[slot][k][m = 4 g + i]
    float p[4];
#pragma unroll
    for (int cc = 0; cc < 4; ++cc) {
      float a = gd[0] * xm[cc][0];
      a = __builtin_fmaf(gd[1], xm[cc][1], a);
      a = __builtin_fmaf(gd[2], xm[cc][2], a);
      p[cc] = __builtin_fmaf(gd[3], xm[cc][3], a);
#pragma unroll
      for (int i = 0; i < 4; ++i) acc2[cc][i] = __builtin_fmaf(gd[i], xkp[cc], acc2[cc][i]);
    }
    // first term of component c = g (channels k < 16 get the second term
    // added at the end)
    reinterpret_cast<float*>(tw)[2 * (4 * sl + g) + half] = row_sums_scatter(p);
  };
  src.issue(0, w1s0, xks0);
  src.issue(1, w1s1, xks1);
  // Step S (S % 6 fixed by the position in the unrolled loop): W1C / XKC
  // stage S, W1N / XKN stage S + 2; XW / XR exchange of steps S / S - 1; TW
  // tile of step S - 1; on odd S >= 3 the tiles T0, T1 of steps S - 3, S - 2
  // are written out first.
#define DPK_L1_BWD_STEP(S, W1C, XKC, W1N, XKN, XW, XR, TW, FLUSH, T0, T1)              \
  if ((S) < nstage) {                                                                  \
    __builtin_amdgcn_s_waitcnt(kWaitVm3Lgkm0);                                         \
    __builtin_amdgcn_s_barrier();                                                      \
    if ((FLUSH) && (S) >= 3) put4(2 * (S) - 6, T0, T1);                                \
    if ((S) > 0) finish(XR, TW);                                                       \
    src.issue((S) + 2, W1N, XKN);                                                      \
    /* fragments of blocks 2 S + u, K chunks kHalfC half + kc: u = 0 at */           \
    /* offsets 0..3 KB, u = 1 at 8..11 KB */                                           \
    const unsigned fa = lds_addr(W1C + (kHalfC * half) * 64 + lane);                  \
    const v4f xkn = lds_read1(lds_addr(XKC + 8 * sl + 4 * half));                     \
    v4f pu[2];                                                                         \
    _Pragma("unroll") for (int u = 0; u < 2; ++u) {                                    \
      v4u af[4];                                                                       \
      if (u == 0)                                                                      \
        lds_read4<0, 1024, 2048, 3072>(af, fa);                                        \
      else                                                                             \
        lds_read4<8192, 9216, 10240, 11264>(af, fa);                                   \
      v4f a = {0.f, 0.f, 0.f, 0.f};                                                    \
      _Pragma("unroll") for (int kc = 0; kc < kHalfC; ++kc) {                          \
        a = E::mfma32(af[kc], gh[kc], a);                                              \
        if constexpr (E::kSplit) a = E::mfma32(af[kc], gl[kc], a);                     \
      }                                                                                \
      pu[u] = a;                                                                       \
    }                                                                                  \
    /* the partner finishes block 2 S + (1 - half) */                                  \
    XW[64 * (wave ^ 1) + lane] = half ? pu[0] : pu[1];                                 \
    own = half ? pu[1] : pu[0];                                                        \
    xkp = xkn;                                                                         \
  }
  for (int s = 0; s < nstage; s += 6) {
    DPK_L1_BWD_STEP(s, w1s0, xks0, w1s2, xks2, xch0, xch1, til2, false, til0, til0)
    DPK_L1_BWD_STEP(s + 1, w1s1, xks1, w1s0, xks0, xch1, xch0, til0, true, til1, til2)
    DPK_L1_BWD_STEP(s + 2, w1s2, xks2, w1s1, xks1, xch0, xch1, til1, false, til0, til0)
    DPK_L1_BWD_STEP(s + 3, w1s0, xks0, w1s2, xks2, xch1, xch0, til2, true, til0, til1)
    DPK_L1_BWD_STEP(s + 4, w1s1, xks1, w1s0, xks0, xch0, xch1, til0, false, til0, til0)
    DPK_L1_BWD_STEP(s + 5, w1s2, xks2, w1s1, xks1, xch1, xch0, til1, true, til2, til0)
  }
#undef DPK_L1_BWD_STEP
  // every copy landed (the ring is reused below) and the last exchange written
  __builtin_amdgcn_s_waitcnt(kWaitVm0Lgkm0);
  __builtin_amdgcn_s_barrier();
  const int last = nstage - 1;
  auto tile = [&](int j) __attribute__((always_inline)) {
    const int r = j % 3;
    return r == 0 ? til0 : (r == 1 ? til1 : til2);
  };
  finish((last & 1) ? xch1 : xch0, tile(last));
  __builtin_amdgcn_s_waitcnt(kWaitVm0Lgkm0);
  __builtin_amdgcn_s_barrier();
  {
    // steps not written out in the loop: from the one after the last pair
    // written (odd S <= last wrote steps S - 3, S - 2) to the last
    const int smax = (last & 1) ? last : last - 1;
    const int first = smax >= 3 ? smax - 1 : 0;
    if (writer) {
      for (int j = first; j <= last; ++j) {
        const float2 v = tile(j)[tix];
        if (2 * j < d3) grow[2 * j] = v.x;
        if (2 * j + 1 < d3) grow[2 * j + 1] = v.y;
      }
    }
  }
  // second term: the pair's partial sums meet in LDS as [group][slot][c][m],
  // half 0 in the first stage buffer, half 1 in the second; the first-term
  // outputs of channels k < 16 are read back (written by other lanes of this
  // workgroup: complete and visible after the wait and barrier)
  __builtin_amdgcn_s_waitcnt(kWaitVm0Lgkm0);
  __builtin_amdgcn_s_barrier();
  float* part0 = reinterpret_cast<float*>(w1s0);
  float* part1 = reinterpret_cast<float*>(w1s1);
  {
    float* mp = (half ? part1 : part0) + ((size_t)grp * 16 + c) * 64;
#pragma unroll
    for (int cc = 0; cc < 4; ++cc)
#pragma unroll
      for (int i = 0; i < 4; ++i) mp[cc * 16 + 4 * g + i] = acc2[cc][i];
  }
  __builtin_amdgcn_s_waitcnt(kWaitVm0Lgkm0);
  __builtin_amdgcn_s_barrier();
  if (valid) {
    float* orow = gxyz + (xa - xyz) + (size_t)g * d3;  // component c = g
    const float* p0 = part0 + ((size_t)grp * 16 + c) * 64 + g * 16;
    const float* p1 = part1 + ((size_t)grp * 16 + c) * 64 + g * 16;
    // channels k < 16 of component g, split between the pair by parity
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int k = 2 * t + half;
      orow[k] += p0[k] + p1[k];
    }
  }
}

}  // namespace l1
}  // namespace
}  // namespace dpk

using namespace dpk;

extern "C" size_t dpk_fit_l1_image_bytes(int d3, int h) {
  if (h != l1::kH || d3 < 16 || d3 > 256) return 0;
  return l1::fwd_bytes(d3) + l1::bwd_bytes(d3);
}

static int l1_check(const char* who, int prec, int d3, int m2, int h) {
  DPK_REQUIRE(prec == DPK_PREC_F16X2, DPK_ERR_UNSUPPORTED,
              "%s: precision %d (the fused layer 1 is built for DPK_PREC_F16X2)", who, prec);
  DPK_REQUIRE(m2 == 16 && h == l1::kH && d3 >= 16 && d3 <= 256, DPK_ERR_UNSUPPORTED,
              "%s: d3=%d m2=%d h=%d (needs m2 = 16, h = 256, 16 <= d3 <= 256)", who, d3, m2, h);
  return DPK_OK;
}

extern "C" int dpk_fit_l1_pack(void* image, int prec, const float* w1, int d3, int m2, int h,
                               dpk_stream_t stream) {
  clear_error();
  const int rc = l1_check("fit_l1_pack", prec, d3, m2, h);
  if (rc != DPK_OK) return rc;
  DPK_REQUIRE(image && w1, DPK_ERR_ARG, "fit_l1_pack: null buffer");
  const long total = (long)dpk_fit_l1_image_bytes(d3, h) / 2;
  hipLaunchKernelGGL(l1::k_pack, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     as_stream(stream), static_cast<unsigned short*>(image), w1, d3);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_fit_l1_fwd(float* z1, int prec, const float* xyz, const int* order, int n,
                              int d3, int m2, const void* image, int h, dpk_stream_t stream) {
  clear_error();
  const int rc = l1_check("fit_l1_fwd", prec, d3, m2, h);
  if (rc != DPK_OK) return rc;
  DPK_REQUIRE(n >= 0, DPK_ERR_ARG, "fit_l1_fwd: n=%d", n);
  DPK_REQUIRE(z1 && xyz && order && image, DPK_ERR_ARG, "fit_l1_fwd: null buffer");
  DPK_REQUIRE((reinterpret_cast<size_t>(z1) & 15) == 0 && (reinterpret_cast<size_t>(image) & 15) == 0,
              DPK_ERR_ARG, "fit_l1_fwd: misaligned buffer");
  if (n == 0) return DPK_OK;
  hipLaunchKernelGGL((l1::k_fwd<h16::F16x2>), dim3(cdiv(n, l1::kSlots)), dim3(l1::kThreads), 0,
                     as_stream(stream), z1, xyz, order,
                     static_cast<const unsigned char*>(image), n, d3);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_fit_l1_bwd(float* gxyz, int prec, const void* gz1, const float* xyz,
                              const int* order, int n, int d3, int m2, const void* image, int h,
                              dpk_stream_t stream) {
  clear_error();
  const int rc = l1_check("fit_l1_bwd", prec, d3, m2, h);
  if (rc != DPK_OK) return rc;
  DPK_REQUIRE(n >= 0, DPK_ERR_ARG, "fit_l1_bwd: n=%d", n);
  DPK_REQUIRE(gxyz && gz1 && xyz && order && image, DPK_ERR_ARG, "fit_l1_bwd: null buffer");
  DPK_REQUIRE((reinterpret_cast<size_t>(gz1) & 15) == 0 && (reinterpret_cast<size_t>(image) & 15) == 0,
              DPK_ERR_ARG, "fit_l1_bwd: misaligned buffer");
  if (n == 0) return DPK_OK;
  const unsigned char* img = static_cast<const unsigned char*>(image) + l1::fwd_bytes(d3);
  hipLaunchKernelGGL((l1::k_bwd<h16::F16x2>), dim3(cdiv(n, l1::kSlots)), dim3(l1::kThreads), 0,
                     as_stream(stream), gxyz,
                     static_cast<const unsigned short*>(gz1), xyz, order, img, n, d3);
  DPK_LAUNCHED();
  return DPK_OK;
}
