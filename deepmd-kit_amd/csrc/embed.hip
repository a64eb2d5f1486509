// Fused se_e2_a embedding net + R^T G contraction on MFMA (forward), and its
// recomputing backward (dE/d em per slot).
//
// Reference math (deepmd/pt/model/descriptor/se_a.py:784-858,
// deepmd/pt/model/network/mlp.py:188-219):
//   per slot j:  s = em[j,0], R = em[j,0:4]
//     y1 = tanh(s w1 + b1)                        (1 -> d1)
//     y2 = tanh(y1 W2 + b2) * idt2 + (y1, y1)     (d1 -> 2 d1, resnet doubling)
//     G  = tanh(y2 W3 + b3) * idt3 + (y2, y2)     (d2 -> 2 d2)
//   xyz = (1/nnei) sum_j R_j^T G_j              [4 x d3]
// Backward: with gx = dE/dxyz / nnei,
//   dE/dR_j[c]  = sum_k gx[c][k] G_j[k]
//   dE/ds_j     = sum_k (sum_c R_j[c] gx[c][k]) dG_j[k]/ds   (tangent, forward mode)
//   net_deriv_j = dE/dR_j + (dE/ds_j, 0, 0, 0)
// Activations are recomputed, never stored (26 M slots x 224 channels at the
// 131k-atom Cu size would be 12 GB in bf16).
//
// MI355X mapping.  One wavefront walks one atom's slots in tiles of 16.  Every
// activation matrix is kept "slot-major on lanes": lane (g = lane/16, c =
// lane%16) holds slot c of the tile and channels 16 m + chan(g, i) (i = 0..3)
// of each 16-channel block m.  That is exactly the C/D layout of a 16x16 MFMA
// whose N index is the slot, and (with the channel block as K) the B-operand
// layout of the next one, so a layer is W^T (A operand, read from the packed
// weights) times the previous activations (B operand, straight from
// registers): no LDS round trip between layers.  The sum over slots of the
// forward contraction needs the slot on the K side: each 16-channel block of
// G is re-laid with channels on lanes and slots on K (through a per-wavefront
// LDS tile for F32/F64; by an identity MFMA, A = G, B = I, in the BF16
// kernels), then R^T x G accumulates xyz.  In the backward, the sums run over
// channels, which are already on K.
//
// Precision: F32 (v_mfma_f32_16x16x4_f32) and F64 (v_mfma_f64_16x16x4_f64)
// run the policy kernels below; BF16 (FP32 accumulate, FP32 activations) has
// its own kernels in namespace bf (LDS-resident weights, K = 32 MFMA).
#include <type_traits>

#include "dpk_common.h"
#include "half16.h"
#include "embed_image.h"

namespace dpk {
namespace {

typedef __attribute__((ext_vector_type(4))) float v4f;
typedef __attribute__((ext_vector_type(4))) double v4d;
typedef __attribute__((ext_vector_type(4))) short v4s;

// 16-bit storage (raw bits) of descriptor / gradient tensors exchanged with
// torch: BF16 or FP16 (h16::bf16s, h16::f16s), round to nearest even
using h16::bf16s;
using h16::f16s;
template <typename T>
__device__ __forceinline__ T to_store(float v) {
  if constexpr (h16::is_h16<T>::value) return T{T::fmt::from_f(v)};
  else return v;
}
__device__ __forceinline__ float from_store(float v) { return v; }
__device__ __forceinline__ double from_store(double v) { return v; }
template <typename T, typename = typename T::fmt>
__device__ __forceinline__ float from_store(T v) { return T::fmt::to_f(v.bits); }

// Branch-free FP32 tanh: 1 - 2 / (2^(2x log2 e) + 1) with the hardware exp2
// and reciprocal (v_exp_f32, v_rcp_f32).  Saturates to +-1 without NaN
// (exp2 -> inf or 0); absolute error ~1e-7, far below the bf16 rounding of the
// GEMM operands and the FP32 path's tolerance.  The library tanhf branches on
// |x| per lane, which diverges inside a wavefront.
__device__ __forceinline__ float fast_tanh(float x) {
  const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// Register i of lane group g of a 16x16 MFMA accumulator holds row chan(g, i)
// (checked on the device by a lane-layout probe): 4g + i for the FP32 and
// BF16 instructions, g + 4i for v_mfma_f64_16x16x4f64.  Every activation
// block uses the same mapping for its channel index, so a layer's output
// registers are the next layer's B operand unchanged.
struct PolF32 {
  using real = float;
  using acc_t = v4f;
  using wt = float;
  static constexpr int kPrec = DPK_PREC_F32;
  __device__ static __forceinline__ int chan(int g, int i) { return 4 * g + i; }
  __device__ static __forceinline__ acc_t zero() { return v4f{0.f, 0.f, 0.f, 0.f}; }
  __device__ static __forceinline__ void mma(acc_t& c, const real a[4], const real b[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[i], c, 0, 0, 0);
  }
  __device__ static __forceinline__ void mma_w(acc_t& c, const wt* aw, int g, const real b[4]) {
    const float4 a4 = *reinterpret_cast<const float4*>(aw + 4 * g);
    const real a[4] = {a4.x, a4.y, a4.z, a4.w};
    mma(c, a, b);
  }
  __device__ static __forceinline__ real tanh_(real x) { return fast_tanh(x); }
};

struct PolF64 {
  using real = double;
  using acc_t = v4d;
  using wt = double;
  static constexpr int kPrec = DPK_PREC_F64;
  __device__ static __forceinline__ int chan(int g, int i) { return g + 4 * i; }
  __device__ static __forceinline__ acc_t zero() { return v4d{0.0, 0.0, 0.0, 0.0}; }
  __device__ static __forceinline__ void mma(acc_t& c, const real a[4], const real b[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[i], c, 0, 0, 0);
  }
  __device__ static __forceinline__ void mma_w(acc_t& c, const wt* aw, int g, const real b[4]) {
    const real a[4] = {aw[g], aw[g + 4], aw[g + 8], aw[g + 12]};
    mma(c, a, b);
  }
  // tanh = 1 - 2 / (e^{2x} + 1): the library exp and a Newton-refined
  // reciprocal (within 1 ulp), absolute error ~2 ulp of 1 and no branches;
  // the library tanh spends several times the instructions (its small-|x|
  // branch keeps the relative error, which the sums of the net do not need).
  // |x| > 20 is clamped (tanh is +-1 there in FP64; keeps e^{2x} finite for
  // the reciprocal's Newton steps); a NaN passes through.
  __device__ static __forceinline__ real tanh_(real x) {
    const real xc = x > 20.0 ? 20.0 : (x < -20.0 ? -20.0 : x);
    return 1.0 - 2.0 * rcp_nr(exp(2.0 * xc) + 1.0);
  }
};

// Compile-time net shape: resnet mode of layers 2 and 3 (0: none, 2: y +=
// (x, x)) and whether the layers carry a time step (idt).
template <int RES2_, int RES3_, bool IDT_>
struct Cfg {
  static constexpr int RES2 = RES2_, RES3 = RES3_;
  static constexpr bool IDT = IDT_;
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Packed per-net parameters (offsets in elements of `real`).
template <int P1, int P2, int P3>
struct NetLayout {
  static constexpr int w1 = 0, b1 = P1;
  static constexpr int w2t = 2 * P1;
  static constexpr int b2 = w2t + P2 * P1, idt2 = b2 + P2;
  static constexpr int w3t = idt2 + P2;
  static constexpr int b3 = w3t + P3 * P2, idt3 = b3 + P3;
  static constexpr int total = idt3 + P3;
};

constexpr int kMaxTypes = 16;
struct BlockInfo {
  int ntypes;
  int sec[kMaxTypes + 1];
  // BF16 kernels only (both may be null): with the formatted neighbour list,
  // a type block is computed up to its last real neighbour; the slots after
  // it (nlist -1, env-matrix rows -avg/std) are constants per (centre type,
  // slot), and pad[((ti * nnei + jj) * 4 + comp) * 128 + internal channel]
  // holds the sum over slots jj..end of jj's block of their R^T G terms.
  const int* nlist;
  const float* pad;
  ActSpec act;  // activation of every layer (the FP32 / FP64 kernels; the others: tanh)
};


template <class Pol, int P1, int P2, int P3>
struct Net {
  using real = typename Pol::real;
  using wt = typename Pol::wt;
  const real* p;   // packed real parameters of this net
  const wt* w2;    // A-operand weights W2^T [P2][P1] (rows s2 apart)
  const wt* w3;    // A-operand weights W3^T [P3][P2] (rows s3 apart)
  int s2 = P1, s3 = P2;
};

template <class Pol, int P1, int P2, int P3>
__device__ __forceinline__ Net<Pol, P1, P2, P3> net_at(const void* params, int k) {
  using L = NetLayout<P1, P2, P3>;
  using real = typename Pol::real;
  Net<Pol, P1, P2, P3> n;
  n.p = static_cast<const real*>(params) + (size_t)k * L::total;
  n.w2 = n.p + L::w2t;
  n.w3 = n.p + L::w3t;
  return n;
}

// LDS-resident weights of the FP64 kernels (one net for the whole launch):
// W2^T and W3^T with rows padded by one element, so that the 16 rows of an
// A fragment (lanes c = 0..15, 16 per LDS pass) fall in distinct banks.
// In FP64 W3^T alone (64 KB) exceeds L1: read from L2 every tile, the
// weights bound the kernels by load latency.
constexpr int kLwWaves = 8;  // atoms (wavefronts) per workgroup sharing the copy
template <int P1, int P2, int P3>
struct LdsW {
  static constexpr int s2 = P1 + 1, s3 = P2 + 1;
  static constexpr int w3 = P2 * s2;
  static constexpr int elems = w3 + P3 * s3;
  // backward: a wavefront's dE/dxyz rows [4][gs] (4 pad elements: the four
  // rows of a read fall in distinct banks)
  static constexpr int gs = P3 + 4;
};

template <class Pol, int P1, int P2, int P3>
__device__ __forceinline__ void stage_weights(typename Pol::real* w, const void* params) {
  using L = NetLayout<P1, P2, P3>;
  using LW = LdsW<P1, P2, P3>;
  using real = typename Pol::real;
  const real* p = static_cast<const real*>(params);
  for (int q = threadIdx.x; q < P2 * P1; q += blockDim.x)
    w[(q / P1) * LW::s2 + q % P1] = p[L::w2t + q];
  for (int q = threadIdx.x; q < P3 * P2; q += blockDim.x)
    w[LW::w3 + (q / P2) * LW::s3 + q % P2] = p[L::w3t + q];
  __syncthreads();
}

template <class Pol, int P1, int P2, int P3>
__device__ __forceinline__ void use_lds_weights(Net<Pol, P1, P2, P3>& n,
                                                const typename Pol::real* w) {
  using LW = LdsW<P1, P2, P3>;
  n.w2 = w;
  n.w3 = w + LW::w3;
  n.s2 = LW::s2;
  n.s3 = LW::s3;
}

// Forward through the three layers for the lane's slot; optionally the
// tangent d/ds.  Each 16-channel block m of the output G (and dG), in the
// slot-on-lane layout, is handed to sink(m, G_m[4], dG_m[4]) as soon as it is
// computed, so the full G never has to live in registers.
// Activation value t and derivative dt at z: tanh (GEN false, the policy's
// tanh) or the net's activation function (GEN true, a wave-uniform switch).
template <class Pol, bool GEN>
__device__ __forceinline__ void act_eval(const ActSpec& a, typename Pol::real z,
                                         typename Pol::real& t, typename Pol::real& dt) {
  using real = typename Pol::real;
  if constexpr (GEN) {
    act_fd<real>(a, z, t, dt);
  } else {
    t = Pol::tanh_(z);
    dt = real(1) - t * t;
  }
}

template <class Pol, int P1, int P2, int P3, class C, bool TANGENT, bool GEN, class Sink>
__device__ __forceinline__ void run_net(const Net<Pol, P1, P2, P3>& net, typename Pol::real s,
                                        int g, int c, const ActSpec& act, Sink&& sink) {
  using real = typename Pol::real;
  using acc_t = typename Pol::acc_t;
  using L = NetLayout<P1, P2, P3>;
  constexpr int M1 = P1 / 16, M2 = P2 / 16, M3 = P3 / 16;
  real y1[M1][4], d1[M1][4];
#pragma unroll
  for (int m = 0; m < M1; ++m)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ch = 16 * m + Pol::chan(g, i);
      const real w = net.p[L::w1 + ch];
      real t, dt;
      act_eval<Pol, GEN>(act, s * w + net.p[L::b1 + ch], t, dt);
      y1[m][i] = t;
      if (TANGENT) d1[m][i] = dt * w;
    }
  real y2[M2][4], d2[M2][4];
#pragma unroll
  for (int m = 0; m < M2; ++m) {
    acc_t a = Pol::zero(), da = Pol::zero();
    const int row = 16 * m + c;
#pragma unroll
    for (int k = 0; k < M1; ++k) {
      Pol::mma_w(a, net.w2 + (size_t)row * net.s2 + 16 * k, g, y1[k]);
      if (TANGENT) Pol::mma_w(da, net.w2 + (size_t)row * net.s2 + 16 * k, g, d1[k]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ch = 16 * m + Pol::chan(g, i);
      real t, dt;
      act_eval<Pol, GEN>(act, a[i] + net.p[L::b2 + ch], t, dt);
      real y = C::IDT ? t * net.p[L::idt2 + ch] : t;
      real dy = real(0);
      if (TANGENT) dy = C::IDT ? dt * da[i] * net.p[L::idt2 + ch] : dt * da[i];
      if constexpr (C::RES2 == 2) {
        y += y1[m % M1][i];
        if (TANGENT) dy += d1[m % M1][i];
      }
      y2[m][i] = y;
      if (TANGENT) d2[m][i] = dy;
    }
    __builtin_amdgcn_sched_barrier(0);
  }
#pragma unroll
  for (int m = 0; m < M3; ++m) {
    acc_t a = Pol::zero(), da = Pol::zero();
    real Gm[4], dGm[4];
    const int row = 16 * m + c;
#pragma unroll
    for (int k = 0; k < M2; ++k) {
      Pol::mma_w(a, net.w3 + (size_t)row * net.s3 + 16 * k, g, y2[k]);
      if (TANGENT) Pol::mma_w(da, net.w3 + (size_t)row * net.s3 + 16 * k, g, d2[k]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ch = 16 * m + Pol::chan(g, i);
      real t, dt;
      act_eval<Pol, GEN>(act, a[i] + net.p[L::b3 + ch], t, dt);
      real y = C::IDT ? t * net.p[L::idt3 + ch] : t;
      real dy = real(0);
      if (TANGENT) dy = C::IDT ? dt * da[i] * net.p[L::idt3 + ch] : dt * da[i];
      if constexpr (C::RES3 == 2) {
        y += y2[m % M2][i];
        if (TANGENT) dy += d2[m % M2][i];
      }
      Gm[i] = y;
      dGm[i] = dy;
    }
    sink(m, Gm, dGm);
    // one 16-channel block at a time: keeps the next block's weight loads
    // from being hoisted (and held live) across this block's work
    __builtin_amdgcn_sched_barrier(0);
  }
}

// --------------------------------------------------------------- forward --
// LW: the launch uses one net, whose W2^T / W3^T the workgroup (kLwWaves
// atoms) holds in LDS (dynamic shared memory); otherwise four atoms per
// workgroup read the weights through the caches.
template <class Pol, int P1, int P2, int P3, class C, bool GEN, bool LW>
__global__ void __launch_bounds__(LW ? 64 * kLwWaves : 256)
    __attribute__((amdgpu_waves_per_eu(Pol::kPrec == DPK_PREC_F64 && !C::IDT ? 2 : 1))) k_embed_fwd(
    typename Pol::real* __restrict__ xyz, const typename Pol::real* __restrict__ em,
    const int* __restrict__ atype, const void* params,
    const int* __restrict__ net_of, const int* __restrict__ ch_out, BlockInfo bi, int d3,
    int natoms, int nnei) {
  using real = typename Pol::real;
  using acc_t = typename Pol::acc_t;
  constexpr int M3 = P3 / 16;
  constexpr int NW = LW ? kLwWaves : 4;
  extern __shared__ __align__(16) unsigned char wsm[];
  real* lw = reinterpret_cast<real*>(wsm);
  if constexpr (LW) stage_weights<Pol, P1, P2, P3>(lw, params);
  const int atom = blockIdx.x * NW + (threadIdx.x >> 6);
  if (atom >= natoms) return;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  int ti = atype[atom];
  ti = ti < 0 ? 0 : ti;
  const real* e = em + (size_t)atom * nnei * 4;
  acc_t acc[M3];
#pragma unroll
  for (int m = 0; m < M3; ++m) acc[m] = Pol::zero();
  // per-wavefront 16 x 16 tile (rows padded to 17) for re-laying G blocks
  __shared__ real tiles[NW][16 * 17];
  real* tile = tiles[threadIdx.x >> 6];
  const bool real_centre = atype[atom] >= 0;
  for (int t = 0; t < bi.ntypes; ++t) {
    Net<Pol, P1, P2, P3> net = net_at<Pol, P1, P2, P3>(params, net_of[ti * bi.ntypes + t]);
    if constexpr (LW) use_lds_weights(net, lw);
    const int b0 = bi.sec[t];
    int b1 = bi.sec[t + 1];
    if (bi.nlist != nullptr) {  // stop at the last real neighbour; padding tail from bi.pad
      b1 = block_end(bi.nlist + (size_t)atom * nnei, b0, b1, lane);
      if (bi.pad != nullptr && b1 < bi.sec[t + 1] && real_centre) {
        const float* pt = bi.pad + ((size_t)ti * nnei + b1) * 4 * P3;
#pragma unroll
        for (int m = 0; m < M3; ++m)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int comp = Pol::chan(g, i);
            if (comp < 4) acc[m][i] += (real)pt[comp * P3 + 16 * m + c];
          }
      }
    }
    for (int s0 = b0; s0 < b1; s0 += 16) {
      // Keep the weights in memory (LDS or the caches), not hoisted into
      // VGPRs for the whole loop: hoisting them costs ~180 VGPRs and drops
      // occupancy to one wave per SIMD.
      asm volatile("" ::: "memory");
      const int slot = s0 + c;
      const real s = slot < b1 ? e[slot * 4] : real(0);
      // R^T operand: row = component c (< 4), K = slots chan(g, i) of the tile
      real rT[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int sl = s0 + Pol::chan(g, i);
        rT[i] = (c < 4 && sl < b1) ? e[sl * 4 + c] : real(0);
      }
      run_net<Pol, P1, P2, P3, C, false, GEN>(net, s, g, c, bi.act,
                                         [&](int m, const real (&Gm)[4], const real (&)[4]) {
        // G block with slots on K: through LDS, row = slot
#pragma unroll
        for (int i = 0; i < 4; ++i) tile[c * 17 + Pol::chan(g, i)] = Gm[i];
        wave_sync();
        real gb[4];  // gb[i] = G[slot chan(g,i)][channel 16m+c]
#pragma unroll
        for (int i = 0; i < 4; ++i) gb[i] = tile[Pol::chan(g, i) * 17 + c];
        wave_sync();
        Pol::mma(acc[m], rT, gb);  // acc[m][i] += sum_slots R[slot][i] G[slot][16m+c]
      });
    }
  }
  // rows 0..3 (the four components) of each accumulator block
  const real inv = real(1) / real(nnei);
  real* o = xyz + (size_t)atom * 4 * d3;
#pragma unroll
  for (int m = 0; m < M3; ++m) {
    const int tc = ch_out[16 * m + c];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int comp = Pol::chan(g, i);
      if (comp < 4 && tc >= 0) o[comp * d3 + tc] = acc[m][i] * inv;
    }
  }
}

// -------------------------------------------------------------- backward --
template <class Pol, int P1, int P2, int P3, class C, bool GEN, bool LW>
__global__ void __launch_bounds__(LW ? 64 * kLwWaves : 256)
    __attribute__((amdgpu_waves_per_eu(Pol::kPrec == DPK_PREC_F64 && !C::IDT ? 2 : 1))) k_embed_bwd(
    typename Pol::real* __restrict__ net_deriv, const typename Pol::real* __restrict__ em,
    const typename Pol::real* __restrict__ gxyz, const int* __restrict__ atype,
    const void* params, const int* __restrict__ net_of,
    const int* __restrict__ ch_out, BlockInfo bi, int d3, int natoms, int nnei) {
  using real = typename Pol::real;
  using acc_t = typename Pol::acc_t;
  constexpr int M3 = P3 / 16;
  constexpr int NW = LW ? kLwWaves : 4;
  extern __shared__ __align__(16) unsigned char wsm[];
  real* lw = reinterpret_cast<real*>(wsm);
  if constexpr (LW) stage_weights<Pol, P1, P2, P3>(lw, params);
  const int atom = blockIdx.x * NW + (threadIdx.x >> 6);
  if (atom >= natoms) return;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  int ti = atype[atom];
  ti = ti < 0 ? 0 : ti;
  const real* e = em + (size_t)atom * nnei * 4;
  real* nd = net_deriv + (size_t)atom * nnei * 4;
  // A operand gx[c][16m + chan(g, i)] / nnei (rows c >= 4 are zero): in
  // registers, or (LW) as the wavefront's four rows [c][128] in LDS, which
  // frees the 64 registers the FP64 backward would otherwise spill
  const real inv = real(1) / real(nnei);
  real gxa[LW ? 1 : M3][4];
  constexpr int GS = LdsW<P1, P2, P3>::gs;
  real* gxl = lw + LdsW<P1, P2, P3>::elems + (size_t)(threadIdx.x >> 6) * 4 * GS;
  if constexpr (LW) {
    for (int q = lane; q < 4 * P3; q += 64) {
      const int tc = ch_out[q % P3];
      gxl[(q / P3) * GS + q % P3] =
          tc >= 0 ? gxyz[(size_t)atom * 4 * d3 + (q / P3) * d3 + tc] * inv : real(0);
    }
    wave_sync();
  } else {
#pragma unroll
    for (int m = 0; m < M3; ++m)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        real v = real(0);
        if (c < 4) {
          const int tc = ch_out[16 * m + Pol::chan(g, i)];
          if (tc >= 0) v = gxyz[(size_t)atom * 4 * d3 + c * d3 + tc] * inv;
        }
        gxa[m][i] = v;
      }
  }
  for (int t = 0; t < bi.ntypes; ++t) {
    Net<Pol, P1, P2, P3> net = net_at<Pol, P1, P2, P3>(params, net_of[ti * bi.ntypes + t]);
    if constexpr (LW) use_lds_weights(net, lw);
    const int b0 = bi.sec[t];
    int b1 = bi.sec[t + 1];
    if (bi.nlist != nullptr) {  // padding slots after the last real neighbour: dE/d em = 0
      b1 = block_end(bi.nlist + (size_t)atom * nnei, b0, b1, lane);
      for (int q = b1 + lane; q < bi.sec[t + 1]; q += 64)
#pragma unroll
        for (int k = 0; k < 4; ++k) nd[(size_t)q * 4 + k] = real(0);
    }
    for (int s0 = b0; s0 < b1; s0 += 16) {
      // Keep the weights in memory (LDS or the caches), not hoisted into
      // VGPRs for the whole loop: hoisting them costs ~180 VGPRs and drops
      // occupancy to one wave per SIMD.
      asm volatile("" ::: "memory");
      const int slot = s0 + c;
      const real s = slot < b1 ? e[slot * 4] : real(0);
      acc_t dR = Pol::zero(), Mx = Pol::zero();
      run_net<Pol, P1, P2, P3, C, true, GEN>(net, s, g, c, bi.act,
                                     [&](int m, const real (&Gm)[4], const real (&dGm)[4]) {
        if constexpr (LW) {
          real ga[4];
#pragma unroll
          for (int i = 0; i < 4; ++i)
            ga[i] = c < 4 ? gxl[c * GS + 16 * m + Pol::chan(g, i)] : real(0);
          Pol::mma(dR, ga, Gm);
          Pol::mma(Mx, ga, dGm);
        } else {
          Pol::mma(dR, gxa[m], Gm);   // dR[comp][slot c], comp = chan(g, i)
          Pol::mma(Mx, gxa[m], dGm);  // sum_k gx[i][k] dG[slot][k]
        }
      });
      // components 0..3 of slot c: (lane c, register comp) for 4g+i
      // layouts, (lane c + 16 comp, register 0) for g+4i
      real dr4[4], mx4[4];
      if constexpr (Pol::kPrec == DPK_PREC_F64) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          dr4[q] = __shfl(dR[0], c + 16 * q, 64);
          mx4[q] = __shfl(Mx[0], c + 16 * q, 64);
        }
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          dr4[q] = dR[q];
          mx4[q] = Mx[q];
        }
      }
      if (g == 0 && slot < b1) {
        const real r0 = e[slot * 4], r1 = e[slot * 4 + 1], r2 = e[slot * 4 + 2],
                   r3 = e[slot * 4 + 3];
        const real ds = r0 * mx4[0] + r1 * mx4[1] + r2 * mx4[2] + r3 * mx4[3];
        real* o = nd + (size_t)slot * 4;
        o[0] = dr4[0] + ds;
        o[1] = dr4[1];
        o[2] = dr4[2];
        o[3] = dr4[3];
      }
    }
  }
}

// ------------------------------------------------------- BF16 fast path --
//
// Same math as the policy kernels above, specialised for BF16 with the
// (32, 64, 128) internal widths:
//  * each net's parameters are staged once per (persistent) workgroup in LDS
//    and shared by its four wavefronts (weights are read every tile);
//  * layers 2 and 3 use v_mfma_f32_16x16x32_bf16 (K = two 16-channel blocks
//    per instruction); the host interleaves each pair of 16-column blocks of
//    W^T so that lane group g's eight K values are one 16-byte LDS read;
//  * FP32 -> BF16 operand packing uses the hardware v_cvt_pk_bf16_f32;
//  * the backward contraction over channels pairs two 16-channel blocks per
//    MFMA.
// The net image (layout in embed_image.h: FP32 vectors, then W2^T and W3^T
// as 16-bit rows in MFMA K order, padded by 8 elements against LDS bank
// conflicts) is what the kernels read, from LDS or global memory.
namespace bf {

using h16::v2u;
using h16::v4u;

using bfimg::P1;
using bfimg::P2;
using bfimg::P3;
using bfimg::W2S;
using bfimg::W3S;
using bfimg::K2L2E;
using bfimg::kInvK2L2E;
using bfimg::NVEC;
using bfimg::V_W1;
using bfimg::V_W1K;
using bfimg::V_B1K;
using bfimg::V_B2K;
using bfimg::V_IDT2;
using bfimg::V_B3K;
using bfimg::V_IDT3;
using bfimg::OFF_W2;
using bfimg::OFF_W3;
using bfimg::NET_BYTES;
using bfimg::tanh_k;
using bfimg::tanh_r;

// Operands are packed pairs of 16-bit values (E = h16::Bf16 or h16::F16),
// each pair converted by one v_cvt_pk_{bf16,f16}_f32 (round to nearest even);
// converting the halves one at a time costs two conversions and a v_perm_b32.
template <class E>
__device__ __forceinline__ v4u pack8(const float a[4], const float b[4]) {
  return v4u{E::pack2(a[0], a[1]), E::pack2(a[2], a[3]), E::pack2(b[0], b[1]),
             E::pack2(b[2], b[3])};
}

template <class E>
__device__ __forceinline__ v2u pack4(float a0, float a1, float a2, float a3) {
  return v2u{E::pack2(a0, a1), E::pack2(a2, a3)};
}

// Low parts of split operands (E::kSplit, FP16): lo = fp16(x - hi) for the
// values packed in hi.
__device__ __forceinline__ unsigned lo_pair(float a, float b, unsigned hi) {
  return h16::F16::pack2(h16::sub_f16_lo(a, hi), h16::sub_f16_hi(b, hi));
}

template <class E>
__device__ __forceinline__ v4u lo8(const float a[4], const float b[4], v4u hi) {
  static_assert(E::kSplit, "split operands are FP16");
  return v4u{lo_pair(a[0], a[1], hi[0]), lo_pair(a[2], a[3], hi[1]), lo_pair(b[0], b[1], hi[2]),
             lo_pair(b[2], b[3], hi[3])};
}

template <class E>
__device__ __forceinline__ v2u lo4(float a0, float a1, float a2, float a3, v2u hi) {
  static_assert(E::kSplit, "split operands are FP16");
  return v2u{lo_pair(a0, a1, hi[0]), lo_pair(a2, a3, hi[1])};
}

__device__ __forceinline__ v4u lds8(const unsigned char* p) {
  return *reinterpret_cast<const v4u*>(p);
}

// cooperative global -> LDS copy of `bytes` (multiple of 16)
__device__ __forceinline__ void stage(unsigned char* dst, const unsigned char* src, int bytes) {
  const int4* s4 = reinterpret_cast<const int4*>(src);
  int4* d4 = reinterpret_cast<int4*>(dst);
  for (int q = threadIdx.x; q < bytes / 16; q += blockDim.x) d4[q] = s4[q];
}



// The net image holds W2, W3 and the biases scaled by K = 2 log2(e), and the
// accumulators start at the scaled bias, so each MFMA returns the tanh
// argument K (W x + b) itself: an activation costs exp, add, rcp and one FMA.
// The tangent (backward) goes through the same scaled weights: its MFMA
// operands are packed as d / K (one multiply per packed value), so the MFMA
// returns the true W d and the tangents stay true derivatives.

// Scaled 16-bit operand of the tangent MFMAs (and its low part when split).
template <class E>
__device__ __forceinline__ v4u pack8_tangent(const float a[4], const float b[4], v4u* lo = nullptr) {
  const float as[4] = {a[0] * kInvK2L2E, a[1] * kInvK2L2E, a[2] * kInvK2L2E, a[3] * kInvK2L2E};
  const float bs[4] = {b[0] * kInvK2L2E, b[1] * kInvK2L2E, b[2] * kInvK2L2E, b[3] * kInvK2L2E};
  const v4u hi = pack8<E>(as, bs);
  if constexpr (E::kSplit) *lo = lo8<E>(as, bs, hi);
  return hi;
}

// Layers 1-2 for the lane's slot: y1 (and dy1/ds) in registers, y2 (and dy2/ds).
// Forward without time steps and with a doubling resnet (FOLD2 / FOLD3), d1
// and d2 carry 1 + y1 and 1 + y2 instead of tangents, so that each resnet
// sum y + tanh is one FMA.
template <class E, bool TANGENT, class C>
__device__ __forceinline__ void layers12(const unsigned char* nb, float s, int g, int c,
                                         float y1[2][4], float d1[2][4],
                                         float y2[4][4], float d2[4][4]) {
  constexpr bool FOLD2 = !TANGENT && !C::IDT && C::RES2 == 2;
  constexpr bool FOLD3 = !TANGENT && !C::IDT && C::RES3 == 2;
  const float* vec = reinterpret_cast<const float*>(nb);
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ch = 16 * m + 4 * g + i;
      const float u = __builtin_fmaf(s, vec[V_W1K + ch], vec[V_B1K + ch]);
      if constexpr (FOLD2) {
        const float r = tanh_r(u);
        y1[m][i] = __builtin_fmaf(-2.0f, r, 1.0f);
        d1[m][i] = __builtin_fmaf(-2.0f, r, 2.0f);
      } else {
        const float t = tanh_k(u);
        y1[m][i] = t;
        if (TANGENT) d1[m][i] = (1.0f - t * t) * vec[V_W1 + ch];
      }
    }
  const v4u B1 = pack8<E>(y1[0], y1[1]);
  v4u B1l, D1, D1l;
  if constexpr (E::kSplit) B1l = lo8<E>(y1[0], y1[1], B1);
  if (TANGENT) D1 = pack8_tangent<E>(d1[0], d1[1], &D1l);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const v4u a = lds8(nb + OFF_W2 + ((16 * m + c) * W2S + 8 * g) * 2);
    v4f acc = E::mfma32(a, B1, *reinterpret_cast<const v4f*>(vec + V_B2K + 16 * m + 4 * g));
    if constexpr (E::kSplit) acc = E::mfma32(a, B1l, acc);
    v4f dacc;
    if (TANGENT) {
      dacc = E::mfma32(a, D1, v4f{0.f, 0.f, 0.f, 0.f});
      if constexpr (E::kSplit) dacc = E::mfma32(a, D1l, dacc);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int ch = 16 * m + 4 * g + i;
      if constexpr (FOLD2) {
        y2[m][i] = __builtin_fmaf(-2.0f, tanh_r(acc[i]), d1[m & 1][i]);
        if constexpr (FOLD3) d2[m][i] = y2[m][i] + 1.0f;
        continue;
      }
      const float t = tanh_k(acc[i]);
      const float idt = C::IDT ? vec[V_IDT2 + ch] : 1.0f;
      float y = C::IDT ? t * idt : t;
      float dy = TANGENT ? (C::IDT ? (1.0f - t * t) * dacc[i] * idt : (1.0f - t * t) * dacc[i])
                         : 0.0f;
      if constexpr (C::RES2 == 2) {
        y += y1[m & 1][i];
        if (TANGENT) dy += d1[m & 1][i];
      }
      y2[m][i] = y;
      if (TANGENT) d2[m][i] = dy;
      if constexpr (FOLD3) d2[m][i] = y + 1.0f;
    }
  }
}

// One 16-channel block m of layer 3: G (and dG/ds) for the lane's slot.
template <class E, bool TANGENT, class C>
__device__ __forceinline__ void layer3_block(const unsigned char* nb, int m, int g, int c,
                                             const v4u B2[2], const v4u D2[2],
                                             const v4u B2l[2], const v4u D2l[2],
                                             const float y2[4][4], const float d2[4][4],
                                             float G[4], float dG[4]) {
  const float* vec = reinterpret_cast<const float*>(nb);
  const unsigned char* row = nb + OFF_W3 + ((16 * m + c) * W3S + 8 * g) * 2;
  const v4u a0 = lds8(row), a1 = lds8(row + 64);
  v4f acc = E::mfma32(a0, B2[0], *reinterpret_cast<const v4f*>(vec + V_B3K + 16 * m + 4 * g));
  acc = E::mfma32(a1, B2[1], acc);
  if constexpr (E::kSplit) {
    acc = E::mfma32(a0, B2l[0], acc);
    acc = E::mfma32(a1, B2l[1], acc);
  }
  v4f dacc;
  if (TANGENT) {
    dacc = E::mfma32(a0, D2[0], v4f{0.f, 0.f, 0.f, 0.f});
    dacc = E::mfma32(a1, D2[1], dacc);
    if constexpr (E::kSplit) {
      dacc = E::mfma32(a0, D2l[0], dacc);
      dacc = E::mfma32(a1, D2l[1], dacc);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int ch = 16 * m + 4 * g + i;
    if constexpr (!TANGENT && !C::IDT && C::RES3 == 2) {  // d2 holds 1 + y2
      G[i] = __builtin_fmaf(-2.0f, tanh_r(acc[i]), d2[m & 3][i]);
      continue;
    }
    const float t = tanh_k(acc[i]);
    const float idt = C::IDT ? vec[V_IDT3 + ch] : 1.0f;
    float y = C::IDT ? t * idt : t;
    float dy = TANGENT ? (C::IDT ? (1.0f - t * t) * dacc[i] * idt : (1.0f - t * t) * dacc[i])
                       : 0.0f;
    // backward with a doubling resnet: the y2 / dy2 terms of G / dG enter
    // the contractions as extra MFMAs on the packed B2 / D2 (k_bwd)
    if constexpr (C::RES3 == 2 && !TANGENT) {
      y += y2[m & 3][i];
    }
    G[i] = y;
    if (TANGENT) dG[i] = dy;
  }
}

// Workgroups (persistent, sized to the occupancy) loop over groups of four
// atoms, one per wavefront.  RESIDENT: every net image is copied to LDS once
// per workgroup; otherwise (many nets) the images are read from global
// memory through the caches.
template <class E, bool RESIDENT, class C>
__global__ void __launch_bounds__(256) k_fwd(
    float* __restrict__ xyz, const float* __restrict__ em, const int* __restrict__ atype,
    const unsigned char* __restrict__ nets, const int* __restrict__ net_of,
    const int* __restrict__ ch_out, BlockInfo bi, int nnets, int d3, int natoms,
    int nnei) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  if (RESIDENT) {
    stage(lds, nets, nnets * NET_BYTES);
    __syncthreads();
  }
  const v2u ident = pack4<E>(c == 4 * g ? 1.f : 0.f, c == 4 * g + 1 ? 1.f : 0.f,
                          c == 4 * g + 2 ? 1.f : 0.f, c == 4 * g + 3 ? 1.f : 0.f);
  for (int base = blockIdx.x * 4; base < natoms; base += gridDim.x * 4) {
    const int atom = base + wave;
    const bool valid = atom < natoms;
    int ti = valid ? atype[atom] : 0;
    ti = ti < 0 ? 0 : ti;
    const float* e = em + (size_t)(valid ? atom : 0) * nnei * 4;
    v4f acc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int t = 0; t < bi.ntypes; ++t) {
      const unsigned char* nb =
          (RESIDENT ? lds : nets) + (size_t)net_of[ti * bi.ntypes + t] * NET_BYTES;
      const int b0 = bi.sec[t];
      int b1 = valid ? bi.sec[t + 1] : b0;
      if (bi.nlist != nullptr && valid) {
        b1 = block_end(bi.nlist + (size_t)atom * nnei, b0, b1, lane);
        // a virtual centre's rows are zero, not padding rows
        if (bi.pad != nullptr && b1 < bi.sec[t + 1] && g == 0 && atype[atom] >= 0) {
          const float* pt = bi.pad + ((size_t)ti * nnei + b1) * 4 * P3;
#pragma unroll
          for (int m = 0; m < 8; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[m][i] += pt[i * P3 + 16 * m + c];
        }
      }
      for (int s0 = b0; s0 < b1; s0 += 16) {
        // weights are re-read from LDS every tile rather than hoisted into
        // registers across the loop (that would halve the occupancy)
        asm volatile("" ::: "memory");
        const int slot = s0 + c;
        const float s = slot < b1 ? e[slot * 4] : 0.f;
        // A operand R^T: rows 0-3 the env-matrix components; split operands
        // put their lo parts in rows 4-7, so one MFMA forms R_hi G and R_lo G
        // (the two row groups are summed once per atom)
        float rr[4];
        constexpr int kRows = E::kSplit ? 8 : 4;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int sl = s0 + 4 * g + i;
          rr[i] = (c < kRows && sl < b1) ? e[sl * 4 + (c & 3)] : 0.f;
        }
        v2u rT = pack4<E>(rr[0], rr[1], rr[2], rr[3]);
        if constexpr (E::kSplit) {
          const v2u rl = lo4<E>(rr[0], rr[1], rr[2], rr[3], rT);
          if (c >= 4) rT = rl;
        }
        float y1[2][4], d1[2][4], y2[4][4], d2[4][4];
        layers12<E, false, C>(nb, s, g, c, y1, d1, y2, d2);
        const v4u B2[2] = {pack8<E>(y2[0], y2[1]), pack8<E>(y2[2], y2[3])};
        v4u B2l[2];
        if constexpr (E::kSplit) {
          B2l[0] = lo8<E>(y2[0], y2[1], B2[0]);
          B2l[1] = lo8<E>(y2[2], y2[3], B2[1]);
        }
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          float G[4], dG[4];
          layer3_block<E, false, C>(nb, m, g, c, B2, B2, B2l, B2l, y2, d2, G, dG);
          // re-lay G with channels on lanes (identity), then R^T G over slots
          const v2u gh = pack4<E>(G[0], G[1], G[2], G[3]);
          const v4f gt = E::mfma16(gh, ident, v4f{0.f, 0.f, 0.f, 0.f});
          const v2u bt = pack4<E>(gt[0], gt[1], gt[2], gt[3]);  // exact: gt holds 16-bit values
          acc[m] = E::mfma16(rT, bt, acc[m]);
          if constexpr (E::kSplit) {
            const v4f gl = E::mfma16(lo4<E>(G[0], G[1], G[2], G[3], gh), ident,
                                     v4f{0.f, 0.f, 0.f, 0.f});
            acc[m] = E::mfma16(rT, pack4<E>(gl[0], gl[1], gl[2], gl[3]), acc[m]);
          }
        }
      }
    }
    if constexpr (E::kSplit) {  // rows 4-7 (lo parts of R) onto rows 0-3
#pragma unroll
      for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[m][i] += __shfl_xor(acc[m][i], 16, 64);
    }
    if (valid && g == 0) {
      const float inv = 1.0f / (float)nnei;
      float* o = xyz + (size_t)atom * 4 * d3;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int tc = ch_out[16 * m + c];
        if (tc >= 0)
#pragma unroll
          for (int i = 0; i < 4; ++i) o[i * d3 + tc] = acc[m][i] * inv;
      }
    }
  }
}

template <class E, bool RESIDENT, class C>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(E::kSplit ? 4 : 5))) k_bwd(
    float* __restrict__ net_deriv, const float* __restrict__ em, const float* __restrict__ gxyz,
    const int* __restrict__ atype, const unsigned char* __restrict__ nets,
    const int* __restrict__ net_of, const int* __restrict__ ch_out, BlockInfo bi,
    int nnets, int d3, int natoms, int nnei) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  if (RESIDENT) {
    stage(lds, nets, nnets * NET_BYTES);
    __syncthreads();
  }
  for (int base = blockIdx.x * 4; base < natoms; base += gridDim.x * 4) {
    const int atom = base + wave;
    const bool valid = atom < natoms;
    int ti = valid ? atype[atom] : 0;
    ti = ti < 0 ? 0 : ti;
    const size_t aoff = (size_t)(valid ? atom : 0);
    const float* e = em + aoff * nnei * 4;
    float* nd = net_deriv + aoff * nnei * 4;
    // A operand gx[c][ch] / nnei for channel-block pairs (rows c >= 4 zero),
    // scaled by a power of two q chosen from the atom's largest |gx| so that
    // the 16-bit operands keep their full significand (dE/dxyz / nnei can sit
    // in FP16's subnormal range); the results are multiplied by 1 / q.
    float amax = 0.f;
    if (valid)  // the atom's 4 d3 values, read by the whole wavefront
      for (int q0 = lane; q0 < 4 * d3; q0 += 64) amax = fmaxf(amax, fabsf(gxyz[aoff * 4 * d3 + q0]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off, 64));
    // every lane holds the maximum: keep q and 1/q in scalar registers
    amax = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(amax)));
    const float q = h16::pow2_scale(amax / (float)nnei);
    const float qin = q / (float)nnei;
    const float unq = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(1.0f / q)));
    // rows 0-3 the components; split operands put their lo parts in rows
    // 4-7, so one MFMA forms gx_hi B and gx_lo B (row groups summed per tile)
    constexpr int kRows = E::kSplit ? 8 : 4;
    const int comp = c & 3;
    v4u gxa[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      float lo[4], hi[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        lo[i] = hi[i] = 0.f;
        if (c < kRows && valid) {
          const int t0 = ch_out[32 * p + 4 * g + i], t1 = ch_out[32 * p + 16 + 4 * g + i];
          if (t0 >= 0) lo[i] = gxyz[aoff * 4 * d3 + comp * d3 + t0] * qin;
          if (t1 >= 0) hi[i] = gxyz[aoff * 4 * d3 + comp * d3 + t1] * qin;
        }
      }
      gxa[p] = pack8<E>(lo, hi);
      if constexpr (E::kSplit) {
        const v4u l = lo8<E>(lo, hi, gxa[p]);
        if (c >= 4) gxa[p] = l;
      }
    }
    for (int t = 0; t < bi.ntypes; ++t) {
      const unsigned char* nb =
          (RESIDENT ? lds : nets) + (size_t)net_of[ti * bi.ntypes + t] * NET_BYTES;
      const int b0 = bi.sec[t];
      int b1 = valid ? bi.sec[t + 1] : b0;
      if (bi.nlist != nullptr && valid) {
        // slots after the last real neighbour move no atom: dE/d em = 0
        b1 = block_end(bi.nlist + (size_t)atom * nnei, b0, b1, lane);
        for (int q = b1 + lane; q < bi.sec[t + 1]; q += 64)
          *reinterpret_cast<float4*>(nd + q * 4) = make_float4(0.f, 0.f, 0.f, 0.f);
      }
      for (int s0 = b0; s0 < b1; s0 += 16) {
        asm volatile("" ::: "memory");
        const int slot = s0 + c;
        const float s = slot < b1 ? e[slot * 4] : 0.f;
        float y1[2][4], d1[2][4], y2[4][4], d2[4][4];
        layers12<E, true, C>(nb, s, g, c, y1, d1, y2, d2);
        const v4u B2[2] = {pack8<E>(y2[0], y2[1]), pack8<E>(y2[2], y2[3])};
        v4u B2l[2], D2l[2];
        const v4u D2[2] = {pack8_tangent<E>(d2[0], d2[1], &D2l[0]),
                           pack8_tangent<E>(d2[2], d2[3], &D2l[1])};
        if constexpr (E::kSplit) {
          B2l[0] = lo8<E>(y2[0], y2[1], B2[0]);
          B2l[1] = lo8<E>(y2[2], y2[3], B2[1]);
        }
        v4f dR = {0.f, 0.f, 0.f, 0.f}, Mx = {0.f, 0.f, 0.f, 0.f}, Mr = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          float G0[4], dG0[4], G1[4], dG1[4];
          layer3_block<E, true, C>(nb, 2 * p, g, c, B2, D2, B2l, D2l, y2, d2, G0, dG0);
          layer3_block<E, true, C>(nb, 2 * p + 1, g, c, B2, D2, B2l, D2l, y2, d2, G1, dG1);
          const v4u gh = pack8<E>(G0, G1), dgh = pack8<E>(dG0, dG1);
          dR = E::mfma32(gxa[p], gh, dR);    // dR[comp][slot c]
          Mx = E::mfma32(gxa[p], dgh, Mx);   // sum_k gx[comp][k] dG[slot][k]
          if constexpr (E::kSplit) {  // (gx_hi, gx_lo) x the lo parts
            dR = E::mfma32(gxa[p], lo8<E>(G0, G1, gh), dR);
            Mx = E::mfma32(gxa[p], lo8<E>(dG0, dG1, dgh), Mx);
          }
          if constexpr (C::RES3 == 2) {
            // resnet copies (y2, y2) and (dy2, dy2): channels 32 p.. of G
            // repeat those of B2 / D2 [p & 1] (D2 holds dy2 / K)
            dR = E::mfma32(gxa[p], B2[p & 1], dR);
            Mr = E::mfma32(gxa[p], D2[p & 1], Mr);
            if constexpr (E::kSplit) {
              dR = E::mfma32(gxa[p], B2l[p & 1], dR);
              Mr = E::mfma32(gxa[p], D2l[p & 1], Mr);
            }
          }
          __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (E::kSplit) {
          // lanes g = 1 hold rows 4-7 (the gx_lo products): fold them into
          // ds and dR of the lanes g = 0 (the components are the same)
          float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
          if (g < 2 && slot < b1) r = *reinterpret_cast<const float4*>(e + slot * 4);
          float ds = r.x * (Mx[0] + K2L2E * Mr[0]) + r.y * (Mx[1] + K2L2E * Mr[1]) +
                     r.z * (Mx[2] + K2L2E * Mr[2]) + r.w * (Mx[3] + K2L2E * Mr[3]);
          ds += __shfl_xor(ds, 16, 64);
#pragma unroll
          for (int i = 0; i < 4; ++i) dR[i] += __shfl_xor(dR[i], 16, 64);
          if (g == 0 && slot < b1)
            *reinterpret_cast<float4*>(nd + slot * 4) =
                make_float4((dR[0] + ds) * unq, dR[1] * unq, dR[2] * unq, dR[3] * unq);
        } else if (g == 0 && slot < b1) {
          const float4 r = *reinterpret_cast<const float4*>(e + slot * 4);
          const float ds = r.x * (Mx[0] + K2L2E * Mr[0]) + r.y * (Mx[1] + K2L2E * Mr[1]) +
                           r.z * (Mx[2] + K2L2E * Mr[2]) + r.w * (Mx[3] + K2L2E * Mr[3]);
          *reinterpret_cast<float4*>(nd + slot * 4) =
              make_float4((dR[0] + ds) * unq, dR[1] * unq, dR[2] * unq, dR[3] * unq);
        }
      }
    }
  }
}

// Net image from the FP32 packed parameters (layout of dpk_embed_net_elems).
// W^T rows are written in MFMA K order: position 32 p + 8 g + j of a row
// holds channel 32 p + 4 g + j (j < 4) or 32 p + 16 + 4 g + (j - 4).
template <class E>
__global__ void k_pack(unsigned char* __restrict__ image, const float* __restrict__ params,
                       int nnets) {
  constexpr float wscale = K2L2E;
  constexpr int ELEMS = 2 * P1 + P2 * P1 + 2 * P2 + P3 * P2 + 2 * P3;
  constexpr int O_W2 = 2 * P1, O_B2 = O_W2 + P2 * P1, O_W3 = O_B2 + 2 * P2, O_B3 = O_W3 + P3 * P2;
  constexpr int NW2 = P2 * W2S, NW3 = P3 * W3S, PER = NVEC + NW2 + NW3;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nnets * PER) return;
  const int k = (int)(q / PER);
  int r = (int)(q % PER);
  const float* p = params + (size_t)k * ELEMS;
  unsigned char* img = image + (size_t)k * NET_BYTES;
  auto kpos = [](int pos) {
    const int pr = pos / 32, w = pos % 32, gg = w / 8, j = w % 8;
    return 32 * pr + (j < 4 ? 4 * gg + j : 16 + 4 * gg + j - 4);
  };
  if (r < NVEC) {
    float v;
    if (r < P1) v = p[r];                                   // w1 (tangent)
    else if (r < 3 * P1) v = p[r - P1] * K2L2E;             // K w1, K b1
    else if (r < 3 * P1 + P2) v = p[O_B2 + r - 3 * P1] * K2L2E;            // K b2
    else if (r < 3 * P1 + 2 * P2) v = p[O_B2 + r - 3 * P1];                // idt2
    else if (r < 3 * P1 + 2 * P2 + P3) v = p[O_B3 + r - 3 * P1 - 2 * P2] * K2L2E;  // K b3
    else v = p[O_B3 + r - 3 * P1 - 2 * P2];                                // idt3
    reinterpret_cast<float*>(img)[r] = v;
    return;
  }
  r -= NVEC;
  if (r < NW2) {
    const int row = r / W2S, pos = r % W2S;
    const float v = pos < P1 ? p[O_W2 + row * P1 + kpos(pos)] * wscale : 0.0f;
    reinterpret_cast<unsigned short*>(img + OFF_W2)[r] = E::from_f(v);
    return;
  }
  r -= NW2;
  const int row = r / W3S, pos = r % W3S;
  const float v = pos < P2 ? p[O_W3 + row * P2 + kpos(pos)] * wscale : 0.0f;
  reinterpret_cast<unsigned short*>(img + OFF_W3)[r] = E::from_f(v);
}

constexpr int kResidentMaxBytes = 64 * 1024;

template <class E, bool RESIDENT, class C>
int launch(bool fwd, void* out, const void* em, const void* gx, const int* atype,
           const void* image, const int* net_of, const int* ch_out, const BlockInfo& bi,
           int nnets, int d3, int natoms, int nnei, hipStream_t st) {
  const size_t lds = RESIDENT ? (size_t)nnets * NET_BYTES : 0;
  // resident workgroups per device, cached per (device, direction, variant, nnets)
  constexpr int kDev = 64, kRes = kResidentMaxBytes / NET_BYTES + 1;
  static int cache[kDev][2][8][kRes];  // one per format E
  int dev = 0;
  DPK_HIP(hipGetDevice(&dev));
  DPK_REQUIRE(dev >= 0 && dev < kDev, DPK_ERR_ARG, "embed: device %d out of range", dev);
  const int variant = (C::RES2 ? 4 : 0) + (C::RES3 ? 2 : 0) + (C::IDT ? 1 : 0);
  int& slots = cache[dev][fwd ? 0 : 1][variant][RESIDENT ? nnets : 0];
  if (slots == 0) {
    int ncu = 0, per_cu = 0;
    DPK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (fwd)
      DPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fwd<E, RESIDENT, C>, 256, lds));
    else
      DPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bwd<E, RESIDENT, C>, 256, lds));
    slots = (ncu < 1 ? 1 : ncu) * (per_cu < 1 ? 1 : per_cu);
  }
  const int groups = cdiv(natoms, 4);
  const int pct = tune(DPK_TUNE_EMBED_GRID_PCT);
  const long want = pct == 0 ? groups : ((long)slots * pct + 99) / 100;
  const int grid = groups < want ? groups : (int)(want > 0 ? want : 1);
  const unsigned char* img = static_cast<const unsigned char*>(image);
  if (fwd)
    hipLaunchKernelGGL((k_fwd<E, RESIDENT, C>), dim3(grid), dim3(256), lds, st,
                       static_cast<float*>(out), static_cast<const float*>(em), atype, img,
                       net_of, ch_out, bi, nnets, d3, natoms, nnei);
  else
    hipLaunchKernelGGL((k_bwd<E, RESIDENT, C>), dim3(grid), dim3(256), lds, st,
                       static_cast<float*>(out), static_cast<const float*>(em),
                       static_cast<const float*>(gx), atype, img, net_of, ch_out, bi, nnets, d3,
                       natoms, nnei);
  DPK_LAUNCHED();
  return DPK_OK;
}

template <class E, bool RESIDENT>
int launch_cfg(bool fwd, void* out, const void* em, const void* gx, const int* atype,
               const void* image, const int* net_of, const int* ch_out, const BlockInfo& bi,
               int res2, int res3, int has_idt, int nnets, int d3, int natoms, int nnei,
               hipStream_t st) {
#define DPK_BF_LAUNCH(R2, R3, I)                                                            \
  return launch<E, RESIDENT, Cfg<R2, R3, I>>(fwd, out, em, gx, atype, image, net_of, ch_out, bi, \
                                          nnets, d3, natoms, nnei, st)
  const int v = (res2 ? 4 : 0) + (res3 ? 2 : 0) + (has_idt ? 1 : 0);
  switch (v) {
    case 0: DPK_BF_LAUNCH(0, 0, false);
    case 1: DPK_BF_LAUNCH(0, 0, true);
    case 2: DPK_BF_LAUNCH(0, 2, false);
    case 3: DPK_BF_LAUNCH(0, 2, true);
    case 4: DPK_BF_LAUNCH(2, 0, false);
    case 5: DPK_BF_LAUNCH(2, 0, true);
    case 6: DPK_BF_LAUNCH(2, 2, false);
    default: DPK_BF_LAUNCH(2, 2, true);
  }
#undef DPK_BF_LAUNCH
}

}  // namespace bf

// ------------------------------------------------------------ descriptor --
template <typename RT, typename OT>
__global__ void __launch_bounds__(256) k_desc_fwd(OT* __restrict__ desc,
                                                  const RT* __restrict__ xyz, int d3, int m2) {
  extern __shared__ __align__(16) unsigned char smem[];
  RT* x = reinterpret_cast<RT*>(smem);
  const int atom = blockIdx.x;
  const RT* xa = xyz + (size_t)atom * 4 * d3;
  for (int q = threadIdx.x; q < 4 * d3; q += blockDim.x) x[q] = xa[q];
  __syncthreads();
  OT* o = desc + (size_t)atom * d3 * m2;
  for (int q = threadIdx.x; q < d3 * m2; q += blockDim.x) {
    const int k = q / m2, m = q % m2;
    RT v = x[k] * x[m] + x[d3 + k] * x[d3 + m] + x[2 * d3 + k] * x[2 * d3 + m] +
           x[3 * d3 + k] * x[3 * d3 + m];
    if constexpr (std::is_same<OT, double>::value) o[q] = v;
    else o[q] = to_store<OT>((float)v);
  }
}

template <typename RT, typename GT>
__global__ void __launch_bounds__(256) k_desc_bwd(RT* __restrict__ gxyz,
                                                  const GT* __restrict__ gdesc,
                                                  const RT* __restrict__ xyz, int d3, int m2) {
  extern __shared__ __align__(16) unsigned char smem[];
  RT* x = reinterpret_cast<RT*>(smem);
  RT* gd = x + 4 * d3;
  const int atom = blockIdx.x;
  const RT* xa = xyz + (size_t)atom * 4 * d3;
  const GT* ga = gdesc + (size_t)atom * d3 * m2;
  for (int q = threadIdx.x; q < 4 * d3; q += blockDim.x) x[q] = xa[q];
  for (int q = threadIdx.x; q < d3 * m2; q += blockDim.x) gd[q] = (RT)from_store(ga[q]);
  __syncthreads();
  RT* o = gxyz + (size_t)atom * 4 * d3;
  for (int q = threadIdx.x; q < 4 * d3; q += blockDim.x) {
    const int cc = q / d3, k = q % d3;
    RT v = RT(0);
    // D[k][m] = sum_c x[c][k] x[c][m]
    for (int m = 0; m < m2; ++m) v += gd[k * m2 + m] * x[cc * d3 + m];
    if (k < m2)
      for (int kk = 0; kk < d3; ++kk) v += gd[kk * m2 + k] * x[cc * d3 + kk];
    o[q] = v;
  }
}

template <int P1, int P2, int P3, class Pol, class C, bool GEN, bool LW>
int launch_embed_w(bool fwd, void* out, const void* em, const void* gx, const int* atype,
                   const void* params, const int* net_of, const int* ch_out, const BlockInfo& bi,
                   int d3, int natoms, int nnei, hipStream_t st) {
  using real = typename Pol::real;
  constexpr int NW = LW ? kLwWaves : 4;
  // LW: the weights, and for the backward each wavefront's dE/dxyz rows
  const size_t lds =
      LW ? sizeof(real) * (LdsW<P1, P2, P3>::elems + (fwd ? 0 : (size_t)NW * 4 * LdsW<P1, P2, P3>::gs))
         : 0;
  dim3 grid(cdiv(natoms, NW)), block(64 * NW);
  if (fwd) {
    hipLaunchKernelGGL((k_embed_fwd<Pol, P1, P2, P3, C, GEN, LW>), grid, block, lds, st,
                       static_cast<real*>(out), static_cast<const real*>(em), atype, params,
                       net_of, ch_out, bi, d3, natoms, nnei);
  } else {
    hipLaunchKernelGGL((k_embed_bwd<Pol, P1, P2, P3, C, GEN, LW>), grid, block, lds, st,
                       static_cast<real*>(out), static_cast<const real*>(em),
                       static_cast<const real*>(gx), atype, params, net_of, ch_out, bi, d3,
                       natoms, nnei);
  }
  DPK_LAUNCHED();
  return DPK_OK;
}

// FP64 with a single net: weights in LDS (a workgroup of kLwWaves atoms);
// FP32 weights (half the bytes) are served by L1 and keep four atoms per
// workgroup.
template <int P1, int P2, int P3, class Pol, class C, bool GEN>
int launch_embed(bool fwd, void* out, const void* em, const void* gx, const int* atype,
                 const void* params, const int* net_of, const int* ch_out, const BlockInfo& bi,
                 int d3, int natoms, int nnei, int nnets, hipStream_t st) {
  if constexpr (Pol::kPrec == DPK_PREC_F64) {
    if (nnets == 1)
      return launch_embed_w<P1, P2, P3, Pol, C, GEN, true>(fwd, out, em, gx, atype, params,
                                                           net_of, ch_out, bi, d3, natoms,
                                                           nnei, st);
  }
  return launch_embed_w<P1, P2, P3, Pol, C, GEN, false>(fwd, out, em, gx, atype, params, net_of,
                                                        ch_out, bi, d3, natoms, nnei, st);
}

template <class Pol>
int launch_embed_cfg(bool fwd, void* out, const void* em, const void* gx, const int* atype,
                     const void* params, const int* net_of, const int* ch_out,
                     const BlockInfo& bi, int res2, int res3, int has_idt, int d3, int natoms,
                     int nnei, int nnets, hipStream_t st) {
#define DPK_POL_LAUNCH(R2, R3, I)                                                            \
  return bi.act.id == DPK_ACT_TANH                                                           \
             ? launch_embed<32, 64, 128, Pol, Cfg<R2, R3, I>, false>(                        \
                   fwd, out, em, gx, atype, params, net_of, ch_out, bi, d3, natoms, nnei,     \
                   nnets, st)                                                                \
             : launch_embed<32, 64, 128, Pol, Cfg<R2, R3, I>, true>(                         \
                   fwd, out, em, gx, atype, params, net_of, ch_out, bi, d3, natoms, nnei,     \
                   nnets, st)
  const int v = (res2 ? 4 : 0) + (res3 ? 2 : 0) + (has_idt ? 1 : 0);
  switch (v) {
    case 0: DPK_POL_LAUNCH(0, 0, false);
    case 1: DPK_POL_LAUNCH(0, 0, true);
    case 2: DPK_POL_LAUNCH(0, 2, false);
    case 3: DPK_POL_LAUNCH(0, 2, true);
    case 4: DPK_POL_LAUNCH(2, 0, false);
    case 5: DPK_POL_LAUNCH(2, 0, true);
    case 6: DPK_POL_LAUNCH(2, 2, false);
    default: DPK_POL_LAUNCH(2, 2, true);
  }
#undef DPK_POL_LAUNCH
}

int embed_common(bool fwd, void* out, int prec, const void* em, const void* gx,
                 const int* atype, const void* params, const void* image, const int* net_of,
                 const int* ch_out, int ntypes, int nnets, int P1, int P2, int P3, int res2,
                 int res3, int has_idt, int act, double act_param, int d3,
                 const int* sec_host, int natoms, int nnei, const int* nlist, const float* pad,
                 dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && nnei > 0, DPK_ERR_ARG, "embed: bad sizes");
  DPK_REQUIRE(ntypes > 0 && ntypes <= kMaxTypes, DPK_ERR_ARG, "embed: ntypes=%d", ntypes);
  DPK_REQUIRE(nnets > 0 && nnets <= ntypes * ntypes, DPK_ERR_ARG, "embed: nnets=%d", nnets);
  DPK_REQUIRE(P1 == 32 && P2 == 64 && P3 == 128, DPK_ERR_UNSUPPORTED,
              "embed: internal widths (%d,%d,%d) unsupported (built for 32,64,128)", P1, P2, P3);
  DPK_REQUIRE((res2 == 0 || res2 == 2) && (res3 == 0 || res3 == 2), DPK_ERR_UNSUPPORTED,
              "embed: resnet modes (%d,%d) unsupported (0 or 2 only)", res2, res3);
  DPK_REQUIRE(d3 > 0 && d3 <= P3, DPK_ERR_ARG, "embed: d3=%d", d3);
  const bool h16prec = prec == DPK_PREC_BF16 || prec == DPK_PREC_F16 || prec == DPK_PREC_F16X2;
  DPK_REQUIRE(h16prec ? image != nullptr : params != nullptr, DPK_ERR_ARG,
              "embed: parameters missing");
  DPK_REQUIRE(pad == nullptr || nlist != nullptr, DPK_ERR_ARG, "embed: pad table without nlist");
  BlockInfo bi;
  DPK_REQUIRE(act_spec(act, act_param, &bi.act), DPK_ERR_ARG, "embed: unknown activation %d", act);
  DPK_REQUIRE(!h16prec || act == DPK_ACT_TANH, DPK_ERR_UNSUPPORTED,
              "embed: the 16-bit kernels are built for tanh (activation %d needs f32 or f64)",
              act);
  bi.ntypes = ntypes;
  bi.nlist = prec != DPK_PREC_F64 ? nlist : nullptr;
  bi.pad = prec != DPK_PREC_F64 ? pad : nullptr;
  for (int t = 0; t <= ntypes; ++t) bi.sec[t] = sec_host[t];
  DPK_REQUIRE(bi.sec[ntypes] == nnei, DPK_ERR_ARG, "embed: sum(sel)=%d != nnei=%d", bi.sec[ntypes], nnei);
  if (natoms == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  switch (prec) {
    case DPK_PREC_BF16: {
      if ((size_t)nnets * bf::NET_BYTES <= (size_t)bf::kResidentMaxBytes)
        return bf::launch_cfg<h16::Bf16, true>(fwd, out, em, gx, atype, image, net_of, ch_out, bi,
                                               res2, res3, has_idt, nnets, d3, natoms, nnei, st);
      return bf::launch_cfg<h16::Bf16, false>(fwd, out, em, gx, atype, image, net_of, ch_out, bi,
                                              res2, res3, has_idt, nnets, d3, natoms, nnei, st);
    }
    case DPK_PREC_F16: {
      if ((size_t)nnets * bf::NET_BYTES <= (size_t)bf::kResidentMaxBytes)
        return bf::launch_cfg<h16::F16, true>(fwd, out, em, gx, atype, image, net_of, ch_out, bi,
                                              res2, res3, has_idt, nnets, d3, natoms, nnei, st);
      return bf::launch_cfg<h16::F16, false>(fwd, out, em, gx, atype, image, net_of, ch_out, bi,
                                             res2, res3, has_idt, nnets, d3, natoms, nnei, st);
    }
    case DPK_PREC_F16X2: {  // the image of dpk_embed_pack_f16 (FP16 weights)
      if ((size_t)nnets * bf::NET_BYTES <= (size_t)bf::kResidentMaxBytes)
        return bf::launch_cfg<h16::F16x2, true>(fwd, out, em, gx, atype, image, net_of, ch_out,
                                                bi, res2, res3, has_idt, nnets, d3, natoms, nnei,
                                                st);
      return bf::launch_cfg<h16::F16x2, false>(fwd, out, em, gx, atype, image, net_of, ch_out, bi,
                                               res2, res3, has_idt, nnets, d3, natoms, nnei, st);
    }
    case DPK_PREC_F32:
      return launch_embed_cfg<PolF32>(fwd, out, em, gx, atype, params, net_of, ch_out, bi, res2,
                                      res3, has_idt, d3, natoms, nnei, nnets, st);
    case DPK_PREC_F64:
      return launch_embed_cfg<PolF64>(fwd, out, em, gx, atype, params, net_of, ch_out, bi, res2,
                                      res3, has_idt, d3, natoms, nnei, nnets, st);
  }
  set_error("embed: unknown precision %d", prec);
  return DPK_ERR_ARG;
}

}  // namespace
}  // namespace dpk

using namespace dpk;

extern "C" size_t dpk_embed_net_elems(int P1, int P2, int P3) {
  return (size_t)2 * P1 + (size_t)P2 * P1 + 2 * (size_t)P2 + (size_t)P3 * P2 + 2 * (size_t)P3;
}

extern "C" size_t dpk_embed_image_bytes(int P1, int P2, int P3) {
  return (P1 == bf::P1 && P2 == bf::P2 && P3 == bf::P3) ? (size_t)bf::NET_BYTES : 0;
}

template <class E>
static int embed_pack(void* image, const float* params, int nnets, int P1, int P2, int P3,
                      dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(P1 == bf::P1 && P2 == bf::P2 && P3 == bf::P3, DPK_ERR_UNSUPPORTED,
              "embed pack: internal widths (%d,%d,%d) unsupported", P1, P2, P3);
  DPK_REQUIRE(nnets >= 0 && image && params, DPK_ERR_ARG, "embed pack: bad arguments");
  if (nnets == 0) return DPK_OK;
  const long n = (long)nnets * (bf::NVEC + bf::P2 * bf::W2S + bf::P3 * bf::W3S);
  hipLaunchKernelGGL(bf::k_pack<E>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     as_stream(stream), static_cast<unsigned char*>(image), params, nnets);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_embed_pack_bf16(void* image, const float* params, int nnets, int P1, int P2,
                                   int P3, dpk_stream_t stream) {
  return embed_pack<h16::Bf16>(image, params, nnets, P1, P2, P3, stream);
}

extern "C" int dpk_embed_pack_f16(void* image, const float* params, int nnets, int P1, int P2,
                                  int P3, dpk_stream_t stream) {
  return embed_pack<h16::F16>(image, params, nnets, P1, P2, P3, stream);
}

extern "C" int dpk_embed_fwd(void* xyz, int prec, const void* em, const int* atype,
                             const void* params, const void* image, const int* net_of,
                             const int* ch_out, int ntypes, int nnets, int P1, int P2, int P3,
                             int res2, int res3, int has_idt, int act, double act_param, int d3,
                             const int* sec_host, int natoms, int nnei, const int* nlist,
                             const float* pad, dpk_stream_t stream) {
  return embed_common(true, xyz, prec, em, nullptr, atype, params, image, net_of, ch_out, ntypes,
                      nnets, P1, P2, P3, res2, res3, has_idt, act, act_param, d3, sec_host, natoms,
                      nnei, nlist, pad, stream);
}

extern "C" int dpk_embed_bwd(void* net_deriv, int prec, const void* em, const void* gxyz,
                             const int* atype, const void* params, const void* image,
                             const int* net_of, const int* ch_out, int ntypes, int nnets, int P1,
                             int P2, int P3, int res2, int res3, int has_idt, int act,
                             double act_param, int d3, const int* sec_host, int natoms, int nnei,
                             const int* nlist, dpk_stream_t stream) {
  return embed_common(false, net_deriv, prec, em, gxyz, atype, params, image, net_of, ch_out,
                      ntypes, nnets, P1, P2, P3, res2, res3, has_idt, act, act_param, d3, sec_host,
                      natoms, nnei, nlist, nullptr, stream);
}

template <typename RT, typename OT>
static int desc_fwd_t(void* desc, const void* xyz, int natoms, int d3, int m2, hipStream_t st) {
  size_t lds = sizeof(RT) * 4 * d3;
  hipLaunchKernelGGL((k_desc_fwd<RT, OT>), dim3(natoms), dim3(256), lds, st,
                     static_cast<OT*>(desc), static_cast<const RT*>(xyz), d3, m2);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_descriptor_fwd(void* desc, int out_prec, const void* xyz, int prec,
                                  int natoms, int d3, int m2, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && d3 > 0 && m2 > 0 && m2 <= d3, DPK_ERR_ARG, "descriptor: bad sizes");
  if (natoms == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  if (prec == DPK_PREC_F64) {
    DPK_REQUIRE(out_prec == DPK_PREC_F64, DPK_ERR_ARG, "descriptor: f64 xyz needs f64 output");
    return desc_fwd_t<double, double>(desc, xyz, natoms, d3, m2, st);
  }
  if (out_prec == DPK_PREC_F32) return desc_fwd_t<float, float>(desc, xyz, natoms, d3, m2, st);
  if (out_prec == DPK_PREC_BF16) return desc_fwd_t<float, bf16s>(desc, xyz, natoms, d3, m2, st);
  if (out_prec == DPK_PREC_F16) return desc_fwd_t<float, f16s>(desc, xyz, natoms, d3, m2, st);
  set_error("descriptor: unsupported output precision %d", out_prec);
  return DPK_ERR_ARG;
}

template <typename RT, typename GT>
static int desc_bwd_t(void* gxyz, const void* gdesc, const void* xyz, int natoms, int d3, int m2,
                      hipStream_t st) {
  size_t lds = sizeof(RT) * (4 * d3 + (size_t)d3 * m2);
  DPK_REQUIRE(lds <= 64 * 1024, DPK_ERR_UNSUPPORTED, "descriptor bwd: LDS %zu too large", lds);
  hipLaunchKernelGGL((k_desc_bwd<RT, GT>), dim3(natoms), dim3(256), lds, st,
                     static_cast<RT*>(gxyz), static_cast<const GT*>(gdesc),
                     static_cast<const RT*>(xyz), d3, m2);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_descriptor_bwd(void* gxyz, int prec, const void* gdesc, int g_prec,
                                  const void* xyz, int natoms, int d3, int m2,
                                  dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && d3 > 0 && m2 > 0 && m2 <= d3, DPK_ERR_ARG, "descriptor: bad sizes");
  if (natoms == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  if (prec == DPK_PREC_F64) {
    DPK_REQUIRE(g_prec == DPK_PREC_F64, DPK_ERR_ARG, "descriptor bwd: f64 needs f64 grad");
    return desc_bwd_t<double, double>(gxyz, gdesc, xyz, natoms, d3, m2, st);
  }
  if (g_prec == DPK_PREC_F32) return desc_bwd_t<float, float>(gxyz, gdesc, xyz, natoms, d3, m2, st);
  if (g_prec == DPK_PREC_BF16)
    return desc_bwd_t<float, bf16s>(gxyz, gdesc, xyz, natoms, d3, m2, st);
  if (g_prec == DPK_PREC_F16) return desc_bwd_t<float, f16s>(gxyz, gdesc, xyz, natoms, d3, m2, st);
  set_error("descriptor bwd: unsupported grad precision %d", g_prec);
  return DPK_ERR_ARG;
}
