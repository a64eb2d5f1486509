// Reverse pass of the se_e2_a embedding net + R^T G contraction for
// training with a force loss: weight gradients of the energy path and of
// the force path (the second derivative), BF16 MFMA with FP32 accumulation.
//
// Reference: the training step differentiates the force -dE/dr with
// create_graph (deepmd/pt/model/model/make_model.py:461,
// transform_output.py:74-86) through DescrptBlockSeA.forward
// (deepmd/pt/model/descriptor/se_a.py:784-858) and the embedding MLPs
// (deepmd/pt/model/network/mlp.py:188-219).  Here the two derivatives are
// two autograd functions over hand-written kernels:
//   X(theta)        : xyz = (1/S) sum_j R_j^T G_j                (embed.hip k_fwd)
//   N(theta, gx)    : nd_j = (1/S)[gx G_j + e0 (R_j^T gx) G'_j]  (embed.hip k_bwd)
// with G' = dG/ds (forward-mode tangent along s = R[0]), and this file's
// kernel for their backward passes:
//   X: given gbar = dL/dxyz, the cotangent on G_j is cp_j = gbar^T R_j / S;
//   N: given v = dL/dnd, cp_j = gx^T v_j / S on G_j and ct_j = v_j[0]
//      (gx^T R_j) / S on G'_j, and dL/dgx = (1/S) sum_j v_j G_j^T +
//      v_j[0] R_j G'_j^T.
// The weight gradient is reverse mode through the net and its tangent
// (t = tanh(z), t' = 1 - t^2, t'' = -2 t t'), per layer with resnet
// doubling y_l = t_l + (y_{l-1}, y_{l-1}):
//   gz_l  = t'_l (a_l - 2 t_l at_l dz_l)      gdz_l = t'_l at_l
//   gW_l += y_{l-1} gz_l^T + dy_{l-1} gdz_l^T  gb_l += gz_l
//   a_{l-1} = W_l gz_l + fold(a_l)            at_{l-1} = W_l gdz_l + fold(at_l)
// (a, at: adjoints of y and of its tangent dy; dz the pre-activation
// tangent), and for layer 1 (z1 = s w1 + b1, dy1 = t'_1 w1):
//   gz1 = t'_1 (a1 - 2 t1 w1 at1), gw1 += s gz1 + at1 t'_1, gb1 += gz1.
// The embedding input s (the env matrix) does not depend on the weights, so
// no gradient with respect to em is formed.
//
// MI355X mapping.  A workgroup of four wavefronts serves one net (grid.y)
// and copies its forward image (W^T, K-scaled, embed_image.h) and backward
// image (W, unscaled) into LDS once.  The four wavefronts take four atoms of
// the net and walk their type block in lockstep tiles of 16 slots; per tile
// each wavefront recomputes its net and tangent in registers (slot on
// lanes, as in embed.hip), runs the adjoints down with W as the A operand
// (output channels on K), and leaves its layer inputs and pre-activation
// gradients in LDS transposed (channel rows, 32 K positions = 16 slots of
// the primal chain then 16 of the tangent chain; transposed by an identity
// MFMA).  Then each wavefront forms a quarter of every weight-gradient
// product over all 64 slots (K = 128) into accumulators it keeps for the
// whole launch; bias gradients come from the same products with a ones
// column.  Each workgroup writes its partial gradient once; a second kernel
// sums the partials in a fixed order (deterministic).

#include "dpk_common.h"
#include "half16.h"
#include "embed_image.h"

namespace dpk {
namespace {
namespace tr {

using h16::v2u;
using h16::v4f;
using h16::v4u;
using E = h16::Bf16;
using namespace bfimg;

// Backward image of one net: W2 [P1 rows][B2S] then W3 [P2 rows][B3S], BF16,
// unscaled; row = input channel, position pos = output channel kpos(pos).
constexpr int B2S = P2 + 8, B3S = P3 + 8;
constexpr int BOFF_W3 = P1 * B2S * 2;
constexpr int BNET_BYTES = BOFF_W3 + P2 * B3S * 2;  // 22016

// Per-wavefront LDS area (BF16 elements): transposed operands of the
// weight-gradient products, rows = channel, KT positions per row, rows KS
// apart (80 bytes: the 16-byte reads of a 16-lane phase, rows c = 0..15,
// then fall in distinct banks; 64-byte rows put rows c and c + 4 together).
constexpr int KT = 32, KS = KT + 8;
constexpr int L_GZ3 = 0;               // [P3][KS]  gz3 | gdz3
constexpr int L_Y2 = L_GZ3 + P3 * KS;  // [P2][KS]  y2 | dy2
constexpr int L_GZ2 = L_Y2 + P2 * KS;  // [P2][KS]  gz2 | gdz2
constexpr int L_Y1 = L_GZ2 + P2 * KS;  // [P1][KS]  y1 | dy1
constexpr int L_L1 = L_Y1 + P1 * KS;   // [P1][KS]  gz1 | at1 t'1
constexpr int L_S = L_L1 + P1 * KS;    // [KT]      s of the slots | 1
constexpr int WAVE_ELEMS = L_S + KT;
constexpr int kWaves = 4;
constexpr int LDS_T = 0;                                   // kWaves areas of WAVE_ELEMS
constexpr int LDS_GM = LDS_T + kWaves * WAVE_ELEMS * 2;    // kWaves x f32 [4][P3]
constexpr int LDS_IMG = LDS_GM + kWaves * 4 * P3 * 4;      // forward image
constexpr int LDS_BIMG = LDS_IMG + NET_BYTES;              // backward image
constexpr int LDS_BYTES = LDS_BIMG + BNET_BYTES;           // 158336
static_assert(LDS_GM % 16 == 0 && LDS_IMG % 16 == 0 && LDS_BIMG % 16 == 0, "LDS alignment");
static_assert(LDS_BYTES <= 160 * 1024, "LDS budget");

constexpr int kMaxNets = 16;
struct NetList {
  int atom_off[kMaxNets + 1];  // atoms of net k: atoms[atom_off[k] .. atom_off[k + 1])
  int b0[kMaxNets], b1[kMaxNets], blk[kMaxNets];  // slot range and type block of net k
};

__device__ __forceinline__ v4u ld16(const unsigned char* p) {
  return *reinterpret_cast<const v4u*>(p);
}

__device__ __forceinline__ v4u pack8(const float a[4], const float b[4]) {
  return v4u{E::pack2(a[0], a[1]), E::pack2(a[2], a[3]), E::pack2(b[0], b[1]),
             E::pack2(b[2], b[3])};
}
__device__ __forceinline__ v2u pack4(const float a[4]) {
  return v2u{E::pack2(a[0], a[1]), E::pack2(a[2], a[3])};
}

// Re-lay a 16x16 block held slot-on-lane (lane (g, c): slot c, channels
// 4g + i) as channel-on-lane (lane (g, c): channel c, slots 4g + i) with an
// identity MFMA; the BF16 values are exact in the FP32 result.
__device__ __forceinline__ v2u relay(const float x[4], v2u ident) {
  const v4f t = E::mfma16(pack4(x), ident, v4f{0.f, 0.f, 0.f, 0.f});
  const float tt[4] = {t[0], t[1], t[2], t[3]};
  return pack4(tt);
}

// Store a re-laid block: row `row0 + c`, K positions koff + 4g .. + 3.
__device__ __forceinline__ void put(unsigned short* area, int row0, int koff, int g, int c,
                                    v2u v) {
  *reinterpret_cast<v2u*>(area + (row0 + c) * KS + koff + 4 * g) = v;
}

__device__ __forceinline__ void copy_lds(unsigned char* dst, const unsigned char* src,
                                         int bytes) {
  const int4* s4 = reinterpret_cast<const int4*>(src);
  int4* d4 = reinterpret_cast<int4*>(dst);
  for (int q = threadIdx.x; q < bytes / 16; q += blockDim.x) d4[q] = s4[q];
}

template <int RES2, int RES3, bool TANG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_vjp(float* __restrict__ parts, float* __restrict__ ggx, const float* __restrict__ em,
      const float* __restrict__ gmat, const float* __restrict__ vnd,
      const unsigned char* __restrict__ image, const unsigned char* __restrict__ bimage,
      const int* __restrict__ atoms, const int* __restrict__ ch_out, NetList nl, int d3,
      int natoms, int nnei) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int k = blockIdx.y;
  copy_lds(lds + LDS_IMG, image + (size_t)k * NET_BYTES, NET_BYTES);
  copy_lds(lds + LDS_BIMG, bimage + (size_t)k * BNET_BYTES, BNET_BYTES);
  const unsigned char* nb = lds + LDS_IMG;
  const unsigned char* bb = lds + LDS_BIMG;
  const float* vec = reinterpret_cast<const float*>(nb);
  unsigned short* tbase = reinterpret_cast<unsigned short*>(lds + LDS_T);
  unsigned short* mine = tbase + wave * WAVE_ELEMS;
  float* gm = reinterpret_cast<float*>(lds + LDS_GM) + wave * 4 * P3;

  const v2u ident = v2u{E::pack2(c == 4 * g ? 1.f : 0.f, c == 4 * g + 1 ? 1.f : 0.f),
                        E::pack2(c == 4 * g + 2 ? 1.f : 0.f, c == 4 * g + 3 ? 1.f : 0.f)};
  // B operand "ones over the primal slots" in column 0 (bias gradients)
  const unsigned one1 = (c == 0 && g < 2) ? E::pack2(1.f, 1.f) : 0u;
  const v4u onesp = v4u{one1, one1, one1, one1};
  const v4u zero4 = v4u{0u, 0u, 0u, 0u};

  v4f acc3[2][5], acc2[3], acc1;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int q = 0; q < 5; ++q) acc3[h][q] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 0; q < 3; ++q) acc2[q] = v4f{0.f, 0.f, 0.f, 0.f};
  acc1 = v4f{0.f, 0.f, 0.f, 0.f};

  const int a0 = nl.atom_off[k], nat = nl.atom_off[k + 1] - a0;
  const int b0 = nl.b0[k], b1 = nl.b1[k], blk = nl.blk[k];
  const int ntile = (b1 - b0 + 15) / 16;
  const float invS = 1.0f / (float)nnei;
  __syncthreads();

  for (int base = blockIdx.x * kWaves; base < nat; base += gridDim.x * kWaves) {
    const int ai = base + wave;
    const bool valid = ai < nat;
    const int atom = valid ? atoms[a0 + ai] : 0;
    // the atom's cotangent matrix / S in internal channel order
    for (int q = lane; q < 4 * P3; q += 64) {
      const int a = q / P3, tc = ch_out[q % P3];
      gm[q] = (valid && tc >= 0) ? gmat[((size_t)atom * 4 + a) * d3 + tc] * invS : 0.f;
    }
    v4f gacc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) gacc[m] = v4f{0.f, 0.f, 0.f, 0.f};
    __syncthreads();

    for (int r = 0; r < ntile; ++r) {
      const int s0 = b0 + 16 * r;
      const int slot = s0 + c;
      const bool sv = valid && slot < b1;
      const size_t srow = (size_t)atom * nnei;
      float4 R = make_float4(0.f, 0.f, 0.f, 0.f), V = R;
      if (sv) {
        R = *reinterpret_cast<const float4*>(em + (srow + slot) * 4);
        if (TANG) V = *reinterpret_cast<const float4*>(vnd + (srow + slot) * 4);
      }
      const float s = R.x;

      // ---- forward: layer 1
      float t1[2][4], dy1[2][4];
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int ch = 16 * m + 4 * g + i;
          const float t = tanh_k(__builtin_fmaf(s, vec[V_W1K + ch], vec[V_B1K + ch]));
          t1[m][i] = t;
          dy1[m][i] = TANG ? (1.0f - t * t) * vec[V_W1 + ch] : 0.f;
        }
      // ---- layer 2 (MFMA from the K-scaled bias: the tanh argument directly)
      const v4u B1 = pack8(t1[0], t1[1]);
      v4u D1 = zero4;
      if (TANG) {
        float a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = dy1[0][i] * kInvK2L2E;
          b[i] = dy1[1][i] * kInvK2L2E;
        }
        D1 = pack8(a, b);
      }
      float t2[4][4], dz2[4][4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const v4u a = ld16(nb + OFF_W2 + ((16 * m + c) * W2S + 8 * g) * 2);
        const v4f acc = E::mfma32(a, B1, *reinterpret_cast<const v4f*>(vec + V_B2K + 16 * m + 4 * g));
        v4f dacc = v4f{0.f, 0.f, 0.f, 0.f};
        if (TANG) dacc = E::mfma32(a, D1, dacc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          t2[m][i] = tanh_k(acc[i]);
          dz2[m][i] = dacc[i];
        }
      }
      float y2[4][4], dy2[4][4];
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          y2[m][i] = t2[m][i] + (RES2 == 2 ? t1[m & 1][i] : 0.f);
          dy2[m][i] = TANG ? (1.0f - t2[m][i] * t2[m][i]) * dz2[m][i] +
                                 (RES2 == 2 ? dy1[m & 1][i] : 0.f)
                           : 0.f;
        }
      const v4u B2[2] = {pack8(y2[0], y2[1]), pack8(y2[2], y2[3])};
      v4u D2[2] = {zero4, zero4};
      if (TANG) {
#pragma unroll
        for (int p = 0; p < 2; ++p) {
          float a[4], b[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            a[i] = dy2[2 * p][i] * kInvK2L2E;
            b[i] = dy2[2 * p + 1][i] * kInvK2L2E;
          }
          D2[p] = pack8(a, b);
        }
      }
      // A operands of the dL/dgx contraction (slots on K): rows 0-3 v_j,
      // and v_j[0] R_j
      v2u Av = v2u{0u, 0u}, Avr = v2u{0u, 0u};
      if (TANG) {
        float va[4], vr[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int sl = s0 + 4 * g + i;
          va[i] = vr[i] = 0.f;
          if (c < 4 && valid && sl < b1) {
            const float vv = vnd[(srow + sl) * 4 + c];
            va[i] = vv;
            vr[i] = vnd[(srow + sl) * 4] * em[(srow + sl) * 4 + c];
          }
        }
        Av = pack4(va);
        Avr = pack4(vr);
      }

      // ---- layer 3, its adjoint, and the layer-3 -> 2 products
      v4f ay2[4], aty2[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) ay2[m] = aty2[m] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        float gz[2][4], gdz[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int m = 2 * p + h;
          const unsigned char* row = nb + OFF_W3 + ((16 * m + c) * W3S + 8 * g) * 2;
          const v4u wa0 = ld16(row), wa1 = ld16(row + 64);
          v4f acc = E::mfma32(wa0, B2[0], *reinterpret_cast<const v4f*>(vec + V_B3K + 16 * m + 4 * g));
          acc = E::mfma32(wa1, B2[1], acc);
          v4f dacc = v4f{0.f, 0.f, 0.f, 0.f};
          if (TANG) {
            dacc = E::mfma32(wa0, D2[0], dacc);
            dacc = E::mfma32(wa1, D2[1], dacc);
          }
          float G[4], dG[4], cp[4], ct[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float t = tanh_k(acc[i]);
            const float tp = 1.0f - t * t;
            G[i] = t + (RES3 == 2 ? y2[m & 3][i] : 0.f);
            dG[i] = TANG ? tp * dacc[i] + (RES3 == 2 ? dy2[m & 3][i] : 0.f) : 0.f;
            const float4 q0 = *reinterpret_cast<const float4*>(gm + 0 * P3 + 16 * m + 4 * g);
            const float4 q1 = *reinterpret_cast<const float4*>(gm + 1 * P3 + 16 * m + 4 * g);
            const float4 q2 = *reinterpret_cast<const float4*>(gm + 2 * P3 + 16 * m + 4 * g);
            const float4 q3 = *reinterpret_cast<const float4*>(gm + 3 * P3 + 16 * m + 4 * g);
            const float g0 = (&q0.x)[i], g1 = (&q1.x)[i], g2 = (&q2.x)[i], g3 = (&q3.x)[i];
            if (TANG) {
              cp[i] = g0 * V.x + g1 * V.y + g2 * V.z + g3 * V.w;
              ct[i] = V.x * (g0 * R.x + g1 * R.y + g2 * R.z + g3 * R.w);
              gz[h][i] = tp * (cp[i] - 2.0f * t * ct[i] * dacc[i]);
              gdz[h][i] = tp * ct[i];
            } else {
              cp[i] = g0 * R.x + g1 * R.y + g2 * R.z + g3 * R.w;
              ct[i] = 0.f;
              gz[h][i] = tp * cp[i];
              gdz[h][i] = 0.f;
            }
          }
          // resnet copies pass the adjoints straight to y2 / dy2
          if (RES3 == 2) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              ay2[m & 3][i] += cp[i];
              if (TANG) aty2[m & 3][i] += ct[i];
            }
          }
          // transposed operands of the layer-3 weight gradient
          put(mine + L_GZ3, 16 * m, 0, g, c, relay(gz[h], ident));
          put(mine + L_GZ3, 16 * m, 16, g, c, TANG ? relay(gdz[h], ident) : v2u{0u, 0u});
          if (TANG) {  // dL/dgx: v^T G + (v0 R)^T G' over the tile's slots
            gacc[m] = E::mfma16(Av, relay(G, ident), gacc[m]);
            gacc[m] = E::mfma16(Avr, relay(dG, ident), gacc[m]);
          }
        }
        // a_y2 += W3 gz3 over this pair's 32 output channels (A = W3 rows)
        const v4u bg = pack8(gz[0], gz[1]);
        const v4u bgd = TANG ? pack8(gdz[0], gdz[1]) : zero4;
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
          const v4u wa = ld16(bb + BOFF_W3 + ((16 * rb + c) * B3S + 32 * p + 8 * g) * 2);
          ay2[rb] = E::mfma32(wa, bg, ay2[rb]);
          if (TANG) aty2[rb] = E::mfma32(wa, bgd, aty2[rb]);
        }
        __builtin_amdgcn_sched_barrier(0);
      }

      // ---- layer 2 adjoint and the layer-2 -> 1 products
      v4f ay1[2], aty1[2];
      ay1[0] = ay1[1] = aty1[0] = aty1[1] = v4f{0.f, 0.f, 0.f, 0.f};
      float gz2[4][4], gdz2[4][4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float t = t2[m][i], tp = 1.0f - t * t;
          const float a = ay2[m][i], at = TANG ? aty2[m][i] : 0.f;
          gz2[m][i] = TANG ? tp * (a - 2.0f * t * at * dz2[m][i]) : tp * a;
          gdz2[m][i] = TANG ? tp * at : 0.f;
          if (RES2 == 2) {
            ay1[m & 1][i] += a;
            if (TANG) aty1[m & 1][i] += at;
          }
        }
        put(mine + L_GZ2, 16 * m, 0, g, c, relay(gz2[m], ident));
        put(mine + L_GZ2, 16 * m, 16, g, c, TANG ? relay(gdz2[m], ident) : v2u{0u, 0u});
        put(mine + L_Y2, 16 * m, 0, g, c, relay(y2[m], ident));
        put(mine + L_Y2, 16 * m, 16, g, c, TANG ? relay(dy2[m], ident) : v2u{0u, 0u});
      }
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const v4u bg = pack8(gz2[2 * p], gz2[2 * p + 1]);
        const v4u bgd = TANG ? pack8(gdz2[2 * p], gdz2[2 * p + 1]) : zero4;
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
          const v4u wa = ld16(bb + ((16 * rb + c) * B2S + 32 * p + 8 * g) * 2);
          ay1[rb] = E::mfma32(wa, bg, ay1[rb]);
          if (TANG) aty1[rb] = E::mfma32(wa, bgd, aty1[rb]);
        }
      }

      // ---- layer 1 adjoint
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        float gz1[4], sec[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int ch = 16 * m + 4 * g + i;
          const float t = t1[m][i], tp = 1.0f - t * t;
          const float a = ay1[m][i], at = TANG ? aty1[m][i] : 0.f;
          gz1[i] = TANG ? tp * (a - 2.0f * t * vec[V_W1 + ch] * at) : tp * a;
          sec[i] = TANG ? tp * at : 0.f;
        }
        put(mine + L_L1, 16 * m, 0, g, c, relay(gz1, ident));
        put(mine + L_L1, 16 * m, 16, g, c, TANG ? relay(sec, ident) : v2u{0u, 0u});
        put(mine + L_Y1, 16 * m, 0, g, c, relay(t1[m], ident));
        put(mine + L_Y1, 16 * m, 16, g, c, TANG ? relay(dy1[m], ident) : v2u{0u, 0u});
      }
      if (g == 0) {
        mine[L_S + c] = E::from_f(s);
        mine[L_S + 16 + c] = E::from_f(1.0f);
      }
      __syncthreads();

      // ---- this wavefront's share of the weight-gradient products, K =
      // the 4 wavefronts x (16 primal + 16 tangent slots)
#pragma unroll
      for (int sw = 0; sw < kWaves; ++sw) {
        const unsigned char* ar = reinterpret_cast<const unsigned char*>(tbase + sw * WAVE_ELEMS);
        v4u A3[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
          A3[h] = ld16(ar + (L_GZ3 + (16 * (2 * wave + h) + c) * KS + 8 * g) * 2);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
          const v4u B = ld16(ar + (L_Y2 + (16 * kb + c) * KS + 8 * g) * 2);
#pragma unroll
          for (int h = 0; h < 2; ++h) acc3[h][kb] = E::mfma32(A3[h], B, acc3[h][kb]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) acc3[h][4] = E::mfma32(A3[h], onesp, acc3[h][4]);
        const v4u A2 = ld16(ar + (L_GZ2 + (16 * wave + c) * KS + 8 * g) * 2);
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
          const v4u B = ld16(ar + (L_Y1 + (16 * kb + c) * KS + 8 * g) * 2);
          acc2[kb] = E::mfma32(A2, B, acc2[kb]);
        }
        acc2[2] = E::mfma32(A2, onesp, acc2[2]);
        if (wave < 2) {
          // column 0: ones over the primal slots (gb1); column 1: s over the
          // primal slots and 1 over the tangent ones (gw1)
          const v4u A1 = ld16(ar + (L_L1 + (16 * wave + c) * KS + 8 * g) * 2);
          const v4u sv = ld16(ar + (L_S + 8 * g) * 2);
          const v4u B = c == 1 ? sv : onesp;
          acc1 = E::mfma32(A1, B, acc1);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      __syncthreads();
    }
    if (TANG && valid && g == 0) {
      float* o = ggx + ((size_t)blk * natoms + atom) * 4 * d3;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const int tc = ch_out[16 * m + c];
        if (tc >= 0)
#pragma unroll
          for (int i = 0; i < 4; ++i) o[i * d3 + tc] = gacc[m][i] * invS;
      }
    }
  }

  // ---- partial gradient of this workgroup (packed parameter layout)
  float* pp = parts + ((size_t)k * gridDim.x + blockIdx.x) * ELEMS;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 16 * (2 * wave + h) + 4 * g + i;
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) pp[E_W3 + o * P2 + 16 * kb + c] = acc3[h][kb][i];
      if (c == 0) pp[E_B3 + o] = acc3[h][4][i];
    }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int o = 16 * wave + 4 * g + i;
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) pp[E_W2 + o * P1 + 16 * kb + c] = acc2[kb][i];
    if (c == 0) pp[E_B2 + o] = acc2[2][i];
  }
  if (wave < 2) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 16 * wave + 4 * g + i;
      if (c == 0) pp[E_B1 + o] = acc1[i];
      if (c == 1) pp[E_W1 + o] = acc1[i];
    }
  }
  // the time steps are not trained by this kernel (nets without idt)
  for (int q = threadIdx.x; q < P2; q += blockDim.x) pp[E_IDT2 + q] = 0.f;
  for (int q = threadIdx.x; q < P3; q += blockDim.x) pp[E_IDT3 + q] = 0.f;
}

// gparams[k][e] = sum over the nx partials of net k in a fixed order: eight
// interleaved partial sums (independent loads in flight), then their sum.
__global__ void k_sum_parts(float* __restrict__ out, const float* __restrict__ parts, int nnets,
                            int nx) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nnets * ELEMS) return;
  const int k = (int)(q / ELEMS), e = (int)(q % ELEMS);
  const float* p = parts + (size_t)k * nx * ELEMS + e;
  float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int x = 0;
  for (; x + 8 <= nx; x += 8)
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += p[(size_t)(x + j) * ELEMS];
  for (int j = 0; x < nx; ++x, ++j) s[j] += p[(size_t)x * ELEMS];
  out[q] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// Backward image from the FP32 packed parameters: W2 / W3 rows by input
// channel, output channels in K order.
__global__ void k_pack_bwd(unsigned char* __restrict__ bimage, const float* __restrict__ params,
                           int nnets) {
  constexpr int N2 = P1 * B2S, N3 = P2 * B3S;
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long)nnets * (N2 + N3)) return;
  const int k = (int)(q / (N2 + N3));
  int r = (int)(q % (N2 + N3));
  const float* p = params + (size_t)k * ELEMS;
  unsigned short* img = reinterpret_cast<unsigned short*>(bimage + (size_t)k * BNET_BYTES);
  if (r < N2) {
    const int row = r / B2S, pos = r % B2S;
    img[r] = E::from_f(pos < P2 ? p[E_W2 + kpos(pos) * P1 + row] : 0.f);
    return;
  }
  r -= N2;
  const int row = r / B3S, pos = r % B3S;
  img[N2 + r] = E::from_f(pos < P3 ? p[E_W3 + kpos(pos) * P2 + row] : 0.f);
}

template <int R2, int R3, bool T>
int launch_vjp(float* parts, float* ggx, const float* em, const float* gmat, const float* vnd,
               const void* image, const void* bimage, const int* atoms, const int* ch_out,
               const NetList& nl, int nnets, int nx, int d3, int natoms, int nnei,
               hipStream_t st) {
  int dev = 0, maxlds = 0;
  DPK_HIP(hipGetDevice(&dev));
  DPK_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  DPK_REQUIRE(maxlds >= LDS_BYTES, DPK_ERR_UNSUPPORTED, "embed train: needs %d bytes of LDS (%d)",
              LDS_BYTES, maxlds);
  hipLaunchKernelGGL((k_vjp<R2, R3, T>), dim3(nx, nnets), dim3(256), LDS_BYTES, st, parts, ggx,
                     em, gmat, vnd, static_cast<const unsigned char*>(image),
                     static_cast<const unsigned char*>(bimage), atoms, ch_out, nl, d3, natoms,
                     nnei);
  DPK_LAUNCHED();
  return DPK_OK;
}

}  // namespace tr
}  // namespace

extern "C" size_t dpk_embed_train_bimage_bytes(void) { return (size_t)tr::BNET_BYTES; }

extern "C" size_t dpk_embed_train_parts_elems(int nnets, int nx) {
  return (size_t)(nnets < 0 ? 0 : nnets) * (size_t)(nx < 0 ? 0 : nx) * bfimg::ELEMS;
}

extern "C" int dpk_embed_train_pack_bwd(void* bimage, const float* params, int nnets,
                                        dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nnets >= 0 && bimage && params, DPK_ERR_ARG, "embed train pack: bad arguments");
  if (nnets == 0) return DPK_OK;
  const long n = (long)nnets * (bfimg::P1 * tr::B2S + bfimg::P2 * tr::B3S);
  hipLaunchKernelGGL(tr::k_pack_bwd, dim3(cdiv(n, 256)), dim3(256), 0, as_stream(stream),
                     static_cast<unsigned char*>(bimage), params, nnets);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_embed_train_vjp(float* gparams, float* ggx, float* parts, int tangent,
                                   const float* em, const float* gmat, const float* vnd,
                                   const void* image, const void* bimage, const int* atoms,
                                   const int* net_atom_off, const int* net_b0,
                                   const int* net_b1, const int* net_blk, const int* ch_out,
                                   int nnets, int nx, int res2, int res3, int has_idt, int d3,
                                   int natoms, int nnei, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nnets >= 1 && nnets <= tr::kMaxNets, DPK_ERR_ARG,
              "embed train: nnets=%d outside 1..%d", nnets, tr::kMaxNets);
  DPK_REQUIRE(nx >= 1, DPK_ERR_ARG, "embed train: nx=%d < 1", nx);
  DPK_REQUIRE(!has_idt, DPK_ERR_UNSUPPORTED,
              "embed train: embedding layers with a time step (resnet_dt) are not supported");
  DPK_REQUIRE((res2 == 0 || res2 == 2) && (res3 == 0 || res3 == 2), DPK_ERR_UNSUPPORTED,
              "embed train: resnet modes (%d,%d) unsupported", res2, res3);
  DPK_REQUIRE(d3 >= 1 && d3 <= bfimg::P3 && natoms >= 0 && nnei >= 1, DPK_ERR_ARG,
              "embed train: bad sizes d3=%d natoms=%d nnei=%d", d3, natoms, nnei);
  DPK_REQUIRE(gparams && parts && em && gmat && image && bimage && atoms && ch_out && net_atom_off &&
                  net_b0 && net_b1 && net_blk,
              DPK_ERR_ARG, "embed train: null argument");
  DPK_REQUIRE(!tangent || (ggx && vnd), DPK_ERR_ARG, "embed train: tangent pass needs ggx and vnd");
  tr::NetList nl{};
  for (int k = 0; k <= nnets; ++k) nl.atom_off[k] = net_atom_off[k];
  for (int k = 0; k < nnets; ++k) {
    DPK_REQUIRE(net_atom_off[k + 1] >= net_atom_off[k], DPK_ERR_ARG,
                "embed train: atom offsets of net %d decrease", k);
    DPK_REQUIRE(net_b0[k] >= 0 && net_b1[k] >= net_b0[k] && net_b1[k] <= nnei, DPK_ERR_ARG,
                "embed train: slot range [%d,%d) of net %d outside [0,%d)", net_b0[k], net_b1[k],
                k, nnei);
    nl.b0[k] = net_b0[k];
    nl.b1[k] = net_b1[k];
    nl.blk[k] = net_blk[k];
  }
  hipStream_t st = as_stream(stream);
#define DPK_TR_LAUNCH(R2, R3, T)                                                                  \
  do {                                                                                          \
    if (int rc = tr::launch_vjp<R2, R3, T>(parts, ggx, em, gmat, vnd, image, bimage, atoms,    \
                                           ch_out, nl, nnets, nx, d3, natoms, nnei, st))       \
      return rc;                                                                                \
  } while (0)
  const int v = (res2 ? 2 : 0) + (res3 ? 1 : 0);
  if (tangent) {
    switch (v) {
      case 0: DPK_TR_LAUNCH(0, 0, true); break;
      case 1: DPK_TR_LAUNCH(0, 2, true); break;
      case 2: DPK_TR_LAUNCH(2, 0, true); break;
      default: DPK_TR_LAUNCH(2, 2, true); break;
    }
  } else {
    switch (v) {
      case 0: DPK_TR_LAUNCH(0, 0, false); break;
      case 1: DPK_TR_LAUNCH(0, 2, false); break;
      case 2: DPK_TR_LAUNCH(2, 0, false); break;
      default: DPK_TR_LAUNCH(2, 2, false); break;
    }
  }
#undef DPK_TR_LAUNCH
  const long n = (long)nnets * bfimg::ELEMS;
  hipLaunchKernelGGL(tr::k_sum_parts, dim3(cdiv(n, 256)), dim3(256), 0, st, gparams, parts, nnets,
                     nx);
  DPK_LAUNCHED();
  return DPK_OK;
}

}  // namespace dpk
