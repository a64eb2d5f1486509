// Layout of the 16-bit embedding-net image read by the MFMA embedding
// kernels (inference in embed.hip, training in embed_train.hip), written on
// the device from the FP32 packed parameters of include/dpk.h.
//
// Internal widths are fixed at (32, 64, 128).  One net's image (bytes):
//   vec f32[NVEC] = w1[32] | K w1[32] | K b1[32] | K b2[64] | idt2[64]
//                   | K b3[128] | idt3[128]          (K = 2 log2 e)
//   W2^T 16-bit [64][W2S]  (row = output channel, K-ordered inputs, scaled by K)
//   W3^T 16-bit [128][W3S]
// Rows are padded by 8 elements against LDS bank conflicts.  Within a row,
// position 32 p + 8 g + j holds input channel kpos(32 p + 8 g + j): the eight
// K values lane group g of a 16x16x32 MFMA reads as one 16-byte load are
// channels 32 p + 4 g + (0..3) and 32 p + 16 + 4 g + (0..3), which are the
// channels that lane group holds in the accumulator layout of two
// consecutive 16-channel blocks.
#pragma once

namespace dpk {
namespace bfimg {

constexpr int P1 = 32, P2 = 64, P3 = 128;
constexpr int W2S = P1 + 8, W3S = P2 + 8;
constexpr float K2L2E = 2.8853900817779268f;
constexpr float kInvK2L2E = 0.34657359027997264f;  // 1 / (2 log2 e) = ln(2) / 2
constexpr int NVEC = 3 * P1 + 2 * P2 + 2 * P3;
constexpr int V_W1 = 0, V_W1K = P1, V_B1K = 2 * P1, V_B2K = 3 * P1, V_IDT2 = 3 * P1 + P2,
              V_B3K = 3 * P1 + 2 * P2, V_IDT3 = 3 * P1 + 2 * P2 + P3;
constexpr int OFF_W2 = NVEC * 4;
constexpr int OFF_W3 = OFF_W2 + P2 * W2S * 2;
constexpr int NET_BYTES = OFF_W3 + P3 * W3S * 2;  // 25472

// FP32 packed parameters of one net (include/dpk.h, dpk_embed_net_elems):
// w1[P1] b1[P1] | w2t[P2][P1] b2[P2] idt2[P2] | w3t[P3][P2] b3[P3] idt3[P3]
constexpr int E_W1 = 0, E_B1 = P1, E_W2 = 2 * P1, E_B2 = E_W2 + P2 * P1, E_IDT2 = E_B2 + P2,
              E_W3 = E_IDT2 + P2, E_B3 = E_W3 + P3 * P2, E_IDT3 = E_B3 + P3,
              ELEMS = E_IDT3 + P3;  // 10688

// Input channel at K position pos of an image row (see above).
__host__ __device__ constexpr int kpos(int pos) {
  return 32 * (pos / 32) +
         ((pos % 32) % 8 < 4 ? 4 * ((pos % 32) / 8) + (pos % 32) % 8
                             : 16 + 4 * ((pos % 32) / 8) + (pos % 32) % 8 - 4);
}

// tanh from the pre-scaled argument u = K z (v_exp_f32 + v_rcp_f32)
__device__ __forceinline__ float tanh_k(float u) {
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(u) + 1.0f);
}
// r with tanh = 1 - 2 r: a resnet sum c + tanh is then fma(-2, r, 1 + c)
__device__ __forceinline__ float tanh_r(float u) {
  return __builtin_amdgcn_rcpf(__builtin_amdgcn_exp2f(u) + 1.0f);
}

}  // namespace bfimg
}  // namespace dpk
