// Fused 16-bit (BF16 or FP16) fitting tail: hidden layers 2..L of width 256, the output layer
// and their backward, for one type segment of slots (atoms grouped by type).
//
// Reference math (deepmd/pt/model/task/fitting.py:734-899,
// deepmd/pt/model/network/network.py:935-1006, mlp.py:188-219):
//   y_1 = tanh(z_1 + b_1) idt_1,                z_1 = D W_1 (the layer-1 GEMM)
//   y_l = tanh(y_{l-1} W_l + b_l) idt_l + res_l y_{l-1}           (l = 2..L)
//   e   = y_L . w_out + ebias
//   g_y_L = w_out,  g_z_l = g_y_l idt_l (1 - t_l^2),
//   g_y_{l-1} = g_z_l W_l^T + res_l g_y_l
// In: z_1 (16-bit, from the layer-1 GEMM).  Out: the atomic energies and g_z_1
// (16-bit), from which the caller forms dE/dD = g_z_1 W_1^T (GEMM).
//
// Products run transposed, Z^T = W^T Y^T, on v_mfma_f32_16x16x32_{bf16,f16} with
// output channels as M, slots as N and input channels as K.  Two wavefronts
// share 16 slots, each owning one half (128) of the channels: lane (g, n)
// holds slot n and channels 16 mb + 4 g + i (i = 0..3) of each of its
// 16-channel blocks mb -- the C layout of the product.  With the K order
// below that is also the B layout of the next product, so activations,
// tanh tangents and gradients stay in registers from z_1 to g_z_1; the two
// halves of each product's B operand are swapped through LDS.  In K chunk kc
// (32 channels) lane group g holds channels 32 kc + 16 (j / 4) + 4 g + j % 4,
// j = 0..7; the weights are packed in that order, one 16-byte A fragment per
// lane per (mb, kc).  A workgroup (8 wavefronts, 64 slots) stages one packed
// 256 x 256 matrix at a time in LDS (128 KB): W_l^T for the forward of layer
// l, W_l for its backward.
#include "dpk_common.h"
#include "half16.h"

namespace dpk {
namespace {
namespace ft {

using h16::v2u;
using h16::v4f;
using h16::v4u;

constexpr int kH = 256;                    // hidden width
constexpr int kMB = kH / 16;               // 16-channel blocks (M)
constexpr int kKC = kH / 32;               // 32-channel K chunks
constexpr int kWaves = 8;
constexpr int kThreads = 64 * kWaves;
constexpr int kTile = 16 * (kWaves / 2);        // slots per workgroup
constexpr int kOwnMB = kMB / 2;                 // blocks owned by a wavefront
constexpr int kOwnKC = kKC / 2;                 // K chunks it produces
constexpr int kWBytes = kMB * kKC * 64 * 16;    // one packed matrix: 128 KB
constexpr int kXBytes = kWaves * kOwnKC * 64 * 16;  // B-operand exchange: 32 KB

// 16-bit formats E = h16::Bf16 or h16::F16: values move as packed pairs
template <class E>
__device__ __forceinline__ v4u pack8(const float a[4], const float b[4]) {
  return v4u{E::pack2(a[0], a[1]), E::pack2(a[2], a[3]), E::pack2(b[0], b[1]),
             E::pack2(b[2], b[3])};
}

// tanh(x) = 1 - 2 / (2^(2 x log2 e) + 1)
__device__ __forceinline__ float tanh_f(float x) {
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(x * 2.8853900817779268f) + 1.0f);
}

// Copy one packed matrix (128 KB) into LDS: 4 passes of 4 16-byte loads per
// thread, each group issued before its stores (a rolled loop, so the
// addresses are not kept live across the kernel).
__device__ __forceinline__ void stage(unsigned char* lds, const unsigned char* src) {
  constexpr int kRounds = kWBytes / 16 / kThreads;
  const int4* s = reinterpret_cast<const int4*>(src) + threadIdx.x;
  int4* d = reinterpret_cast<int4*>(lds) + threadIdx.x;
#pragma unroll 1
  for (int r0 = 0; r0 < kRounds; r0 += 4) {
    int4 v[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = s[(r0 + r) * kThreads];
#pragma unroll
    for (int r = 0; r < 4; ++r) d[(r0 + r) * kThreads] = v[r];
  }
}

// B-operand halves through LDS: wavefront w writes its chunks, its partner
// w ^ 1 reads them (same lane, same chunk position).
__device__ __forceinline__ void xwrite(unsigned char* xb, int wave, int lane, const v4u own[kOwnKC]) {
#pragma unroll
  for (int k = 0; k < kOwnKC; ++k)
    *reinterpret_cast<v4u*>(xb + (((wave * kOwnKC + k) * 64 + lane) << 4)) = own[k];
}
__device__ __forceinline__ void xread(const unsigned char* xb, int wave, int lane, v4u oth[kOwnKC]) {
#pragma unroll
  for (int k = 0; k < kOwnKC; ++k)
    oth[k] = *reinterpret_cast<const v4u*>(xb + ((((wave ^ 1) * kOwnKC + k) * 64 + lane) << 4));
}

// Two of the wavefront's blocks (b0, b0 + 1) of one product: acc += W-rows x B.
// Fragment (mb, kc) of a packed matrix sits at ((mb * 8 + kc) * 64 + lane) * 16;
// for the wavefront's blocks mb = 8 h + b and its own chunks kc = 4 h + k that
// is fo + (8 b + k) * 1024, for the partner's chunks kc = 4 (1 - h) + k it is
// fp + (8 b + k) * 1024 -- compile-time LDS offsets on two base addresses.
template <class E>
__device__ __forceinline__ void product2(const unsigned char* fo, const unsigned char* fp, int b0,
                                         const v4u own[kOwnKC], const v4u oth[kOwnKC], v4f& a0,
                                         v4f& a1) {
#pragma unroll
  for (int k = 0; k < kOwnKC; ++k) {
    a0 = E::mfma32(*reinterpret_cast<const v4u*>(fo + (8 * b0 + k) * 1024), own[k], a0);
    a1 = E::mfma32(*reinterpret_cast<const v4u*>(fo + (8 * b0 + 8 + k) * 1024), own[k], a1);
  }
  asm volatile("" ::: "memory");  // at most 8 fragments in flight
#pragma unroll
  for (int k = 0; k < kOwnKC; ++k) {
    a0 = E::mfma32(*reinterpret_cast<const v4u*>(fp + (8 * b0 + k) * 1024), oth[k], a0);
    a1 = E::mfma32(*reinterpret_cast<const v4u*>(fp + (8 * b0 + 8 + k) * 1024), oth[k], a1);
  }
}

template <class E>
__device__ __forceinline__ void pack_own(const float y[kOwnMB][4], v4u own[kOwnKC]) {
#pragma unroll
  for (int k = 0; k < kOwnKC; ++k) own[k] = pack8<E>(y[2 * k], y[2 * k + 1]);
}

// own gradient chunks g_y * s (s as 16-bit pairs)
template <class E>
__device__ __forceinline__ void pack_grad(const float gy[kOwnMB][4], const unsigned s[2 * kOwnMB],
                                          v4u own[kOwnKC]) {
#pragma unroll
  for (int k = 0; k < kOwnKC; ++k) {
    const unsigned* s0 = &s[4 * k];
    const float a[4] = {gy[2 * k][0] * E::lo(s0[0]), gy[2 * k][1] * E::hi(s0[0]),
                        gy[2 * k][2] * E::lo(s0[1]), gy[2 * k][3] * E::hi(s0[1])};
    const float b[4] = {gy[2 * k + 1][0] * E::lo(s0[2]), gy[2 * k + 1][1] * E::hi(s0[2]),
                        gy[2 * k + 1][2] * E::lo(s0[3]), gy[2 * k + 1][3] * E::hi(s0[3])};
    own[k] = pack8<E>(a, b);
  }
}

// Image of one net (dpk_fit_tail_bytes): the L-1 forward matrices W_l^T, the
// L-1 backward matrices W_l (packed), then FP32 vectors of 256: b_1, idt_1,
// (b_l, idt_l) for l = 2..L, w_out.
// z_1 of the lane's four channels i of block b (16-bit pairs, or FP32 in the
// split mode, where z_1 comes out of the GEMM unrounded)
template <class E>
__device__ __forceinline__ void load_z4(const void* zr, int b, float zf[4]) {
  if constexpr (E::kSplit) {
    const v4f z = *reinterpret_cast<const v4f*>(static_cast<const float*>(zr) + 16 * b);
    zf[0] = z[0]; zf[1] = z[1]; zf[2] = z[2]; zf[3] = z[3];
  } else {
    const v2u zz = *reinterpret_cast<const v2u*>(static_cast<const unsigned short*>(zr) + 16 * b);
    zf[0] = E::lo(zz[0]); zf[1] = E::hi(zz[0]); zf[2] = E::lo(zz[1]); zf[3] = E::hi(zz[1]);
  }
}

template <class E, int L>
__global__ void __launch_bounds__(kThreads) k_fit_tail(double* __restrict__ atom_e,
                                                       unsigned short* __restrict__ gz1,
                                                       const void* __restrict__ z1,
                                                       const unsigned char* __restrict__ net,
                                                       int res_mask, double ebias,
                                                       const int* __restrict__ order, int n) {
  extern __shared__ __align__(16) unsigned char lds[];
  unsigned char* xb = lds + kWBytes;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane >> 4, c = lane & 15;
  const int h = wave & 1;  // channel half
  const int slot = blockIdx.x * kTile + 16 * (wave >> 1) + c;
  const bool valid = slot < n;
  const int cb = 128 * h + 4 * g;  // channel of block 0, i = 0; block b adds 16 b
  const float* vec = reinterpret_cast<const float*>(net + (size_t)2 * (L - 1) * kWBytes);
  const float* wout = vec + 2 * L * kH;
  const unsigned char* fo = lds + ((68 * h) * 64 + lane) * 16;
  const unsigned char* fp = lds + ((60 * h + 4) * 64 + lane) * 16;

  stage(lds, net);  // W_2^T

  // layer 1 epilogue: y_1 (s_1 is recomputed from z_1 at the end)
  float y[kOwnMB][4];
  unsigned sp[L][2 * kOwnMB];  // s_l = idt_l (1 - t_l^2) as 16-bit pairs, l = 2..L
  v4u own[kOwnKC], oth[kOwnKC];
  const size_t zoff = (size_t)(valid ? slot : 0) * kH + cb;
  const void* zr = E::kSplit ? (const void*)(static_cast<const float*>(z1) + zoff)
                             : (const void*)(static_cast<const unsigned short*>(z1) + zoff);
#pragma unroll
  for (int b = 0; b < kOwnMB; ++b) {
    float zf[4] = {0.f, 0.f, 0.f, 0.f};
    if (valid) load_z4<E>(zr, b, zf);
    const v4f bb = *reinterpret_cast<const v4f*>(vec + cb + 16 * b);
    const v4f ii = *reinterpret_cast<const v4f*>(vec + kH + cb + 16 * b);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[b][i] = tanh_f(zf[i] + bb[i]) * ii[i];
  }
  pack_own<E>(y, own);
  xwrite(xb, wave, lane, own);
  __syncthreads();  // W_2^T and the exchange written
  xread(xb, wave, lane, oth);

  // forward through layers 2..L
#pragma unroll
  for (int t = 0; t < L - 1; ++t) {
    const float* bl = vec + (2 + 2 * t) * kH + cb;
    const float* il = vec + (3 + 2 * t) * kH + cb;
    const bool res = (res_mask >> t) & 1;
#pragma unroll
    for (int b = 0; b < kOwnMB; b += 2) {
      v4f a0 = *reinterpret_cast<const v4f*>(bl + 16 * b);
      v4f a1 = *reinterpret_cast<const v4f*>(bl + 16 * b + 16);
      product2<E>(fo, fp, b, own, oth, a0, a1);
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const v4f acc = u ? a1 : a0;
        const v4f ii = *reinterpret_cast<const v4f*>(il + 16 * (b + u));
        float s[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float th = tanh_f(acc[i]);
          y[b + u][i] = th * ii[i] + (res ? y[b + u][i] : 0.0f);
          s[i] = ii[i] * (1.0f - th * th);
        }
        sp[t + 1][2 * (b + u)] = E::pack2(s[0], s[1]);
        sp[t + 1][2 * (b + u) + 1] = E::pack2(s[2], s[3]);
      }
      asm volatile("" ::: "memory");  // keep the next blocks' loads here
    }
    if (t < L - 2) {
      pack_own<E>(y, own);
      __syncthreads();  // every wavefront is done with the matrix and the exchange
      xwrite(xb, wave, lane, own);
      stage(lds, net + (size_t)(t + 1) * kWBytes);
      __syncthreads();
      xread(xb, wave, lane, oth);
    }
  }

  // output layer: e = y_L . w_out; lanes n, n+16, n+32, n+48 share slot n and
  // the two halves meet in atom_e (half 1 stores its part before the barrier)
  float gy[kOwnMB][4];
  float e = 0.0f;
#pragma unroll
  for (int b = 0; b < kOwnMB; ++b) {
    const v4f w = *reinterpret_cast<const v4f*>(wout + cb + 16 * b);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      e = __builtin_fmaf(y[b][i], w[i], e);
      gy[b][i] = w[i];
    }
  }
  e += __shfl_xor(e, 16, 64);
  e += __shfl_xor(e, 32, 64);
  if (h == 1 && g == 0 && valid) atom_e[order[slot]] = (double)e;
  pack_grad<E>(gy, sp[L - 1], own);
  __syncthreads();
  if (h == 0 && g == 0 && valid) atom_e[order[slot]] += (double)e + ebias;
  xwrite(xb, wave, lane, own);
  stage(lds, net + (size_t)(L - 1 + L - 2) * kWBytes);  // W_L
  __syncthreads();
  xread(xb, wave, lane, oth);

  // backward through layers L..2: g_y_{l-1} = (g_y_l * s_l) W_l^T + res_l g_y_l
#pragma unroll
  for (int t = L - 2; t >= 0; --t) {
    const bool res = (res_mask >> t) & 1;
#pragma unroll
    for (int b = 0; b < kOwnMB; b += 2) {
      v4f a0 = res ? v4f{gy[b][0], gy[b][1], gy[b][2], gy[b][3]} : v4f{0.f, 0.f, 0.f, 0.f};
      v4f a1 = res ? v4f{gy[b + 1][0], gy[b + 1][1], gy[b + 1][2], gy[b + 1][3]}
                   : v4f{0.f, 0.f, 0.f, 0.f};
      product2<E>(fo, fp, b, own, oth, a0, a1);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        gy[b][i] = a0[i];
        gy[b + 1][i] = a1[i];
      }
      asm volatile("" ::: "memory");
    }
    if (t > 0) {
      pack_grad<E>(gy, sp[t], own);
      __syncthreads();
      xwrite(xb, wave, lane, own);
      stage(lds, net + (size_t)(L - 1 + t - 1) * kWBytes);  // W_{t+1}
      __syncthreads();
      xread(xb, wave, lane, oth);
    }
  }

  // g_z_1 = g_y_1 * s_1 (16-bit), s_1 = idt_1 (1 - t_1^2) rounded to 16 bits as
  // the other layers' s
  // (split mode: s_1 in FP32 and g_z_1 as hi / lo halves of a 2 x 256 row)
  if (valid) {
    unsigned short* gr = gz1 + (size_t)slot * (E::kSplit ? 2 * kH : kH) + cb;
#pragma unroll
    for (int b = 0; b < kOwnMB; ++b) {
      float zf[4];
      load_z4<E>(zr, b, zf);
      const v4f bb = *reinterpret_cast<const v4f*>(vec + cb + 16 * b);
      const v4f ii = *reinterpret_cast<const v4f*>(vec + kH + cb + 16 * b);
      float s[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float t = tanh_f(zf[i] + bb[i]);
        s[i] = ii[i] * (1.0f - t * t);
      }
      if constexpr (E::kSplit) {
        float gzv[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) gzv[i] = gy[b][i] * s[i];
        const v2u h = {E::pack2(gzv[0], gzv[1]), E::pack2(gzv[2], gzv[3])};
        const v2u l = {E::pack2(gzv[0] - E::lo(h[0]), gzv[1] - E::hi(h[0])),
                       E::pack2(gzv[2] - E::lo(h[1]), gzv[3] - E::hi(h[1]))};
        *reinterpret_cast<v2u*>(gr + 16 * b) = h;
        *reinterpret_cast<v2u*>(gr + kH + 16 * b) = l;
      } else {
        const unsigned s01 = E::pack2(s[0], s[1]), s23 = E::pack2(s[2], s[3]);
        const v2u o = {E::pack2(gy[b][0] * E::lo(s01), gy[b][1] * E::hi(s01)),
                       E::pack2(gy[b][2] * E::lo(s23), gy[b][3] * E::hi(s23))};
        *reinterpret_cast<v2u*>(gr + 16 * b) = o;
      }
    }
  }
}

}  // namespace ft
}  // namespace
}  // namespace dpk

using namespace dpk;

extern "C" size_t dpk_fit_tail_bytes(int L) {
  if (L < 2 || L > 3) return 0;
  return (size_t)2 * (L - 1) * ft::kWBytes + sizeof(float) * (size_t)(2 * L + 1) * ft::kH;
}

template <class E>
static int fit_tail(double* atom_energy, void* gz1, const void* z1, const void* net, int L,
                    int res_mask, double ebias, const int* order, int n, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(L >= 2 && L <= 3, DPK_ERR_UNSUPPORTED, "fit_tail: %d hidden layers (2 or 3)", L);
  DPK_REQUIRE(n >= 0, DPK_ERR_ARG, "fit_tail: n=%d", n);
  DPK_REQUIRE(atom_energy && gz1 && z1 && net && order, DPK_ERR_ARG, "fit_tail: null buffer");
  DPK_REQUIRE((reinterpret_cast<size_t>(net) & 15) == 0 &&
                  (reinterpret_cast<size_t>(z1) & (E::kSplit ? 15 : 7)) == 0 &&
                  (reinterpret_cast<size_t>(gz1) & 7) == 0,
              DPK_ERR_ARG, "fit_tail: misaligned buffer");
  if (n == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  int dev = 0, maxlds = 0;
  DPK_HIP(hipGetDevice(&dev));
  DPK_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  constexpr int kLds = ft::kWBytes + ft::kXBytes;
  DPK_REQUIRE(maxlds >= kLds, DPK_ERR_UNSUPPORTED, "fit_tail: needs %d bytes of LDS (%d)", kLds,
              maxlds);
  const dim3 grid(cdiv(n, ft::kTile)), block(ft::kThreads);
  auto* ae = atom_energy;
  auto* gz = static_cast<unsigned short*>(gz1);
  const void* zz = z1;
  auto* nb = static_cast<const unsigned char*>(net);
  if (L == 2)
    hipLaunchKernelGGL((ft::k_fit_tail<E, 2>), grid, block, kLds, st, ae, gz, zz, nb, res_mask,
                       ebias, order, n);
  else
    hipLaunchKernelGGL((ft::k_fit_tail<E, 3>), grid, block, kLds, st, ae, gz, zz, nb, res_mask,
                       ebias, order, n);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_fit_tail_bf16(double* atom_energy, void* gz1, const void* z1, const void* net,
                                 int L, int res_mask, double ebias, const int* order, int n,
                                 dpk_stream_t stream) {
  return fit_tail<h16::Bf16>(atom_energy, gz1, z1, net, L, res_mask, ebias, order, n, stream);
}

extern "C" int dpk_fit_tail_f16(double* atom_energy, void* gz1, const void* z1, const void* net,
                                int L, int res_mask, double ebias, const int* order, int n,
                                dpk_stream_t stream) {
  return fit_tail<h16::F16>(atom_energy, gz1, z1, net, L, res_mask, ebias, order, n, stream);
}

extern "C" int dpk_fit_tail_f16x2(double* atom_energy, void* gz1, const void* z1,
                                  const void* net, int L, int res_mask, double ebias,
                                  const int* order, int n, dpk_stream_t stream) {
  return fit_tail<h16::F16x2>(atom_energy, gz1, z1, net, L, res_mask, ebias, order, n, stream);
}
