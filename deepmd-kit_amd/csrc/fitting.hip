// Descriptor and energy-fitting pieces of the se_e2_a path around the
// fitting GEMMs (which run on hipBLASLt).
//
// Reference math (deepmd/pt/model/descriptor/se_a.py:844-850,
// deepmd/pt/model/task/fitting.py:734-899, deepmd/pt/model/network/mlp.py:
// 188-219):
//   D[k][m] = sum_c xyz[c][k] xyz[c][m]          (k < d3, m < M2), row k*M2+m
//   hidden layer l:  y_l = tanh(z_l + b_l) * idt_l + res_l(y_{l-1}),
//                    z_l = y_{l-1} W_l           (GEMM)
//   output:          e = y_L . w_out + b_out + bias_atom_e[type]
//   backward:        g_y_L = w_out, g_z_l = g_y_l * idt_l (1 - t_l^2),
//                    g_y_{l-1} = g_z_l W_l^T (GEMM) + res_l^T(g_y_l),
//                    dE/dD = g_z_1 W_1^T (GEMM)
//   dE/dxyz[c][p] = sum_m dE/dD[p][m] xyz[c][m] + [p < M2] sum_k dE/dD[k][p] xyz[c][k]
// Rows are "slots": atoms grouped by type (slot i is atom order[i]) so that
// each type's fitting net is one contiguous GEMM.  res is 0 (none), 1 (y +=
// x) or 2 (y += (x, x)).  Every kernel is elementwise or one wavefront per
// row and HBM-bound.
#include <type_traits>

#include "dpk_common.h"
#include "half16.h"

namespace dpk {
namespace {
namespace fit {

// 16-bit GEMM operand storage: BF16 or FP16 bits (h16::bf16s / h16::f16s)
using h16::bf16s;
using h16::f16s;
using h16::is_h16;
// Split FP16 descriptor rows (DPK_PREC_F16X2): a row of 2 d3 M2 halves, the
// hi parts fp16(D) then the lo parts fp16(D - hi), so that one GEMM with
// K = 2 d3 M2 against the stacked [W; W] forms D W to ~22 bits.
struct f16x2s {
  unsigned short bits;
};

template <typename T>
__device__ __forceinline__ float to_f(T v) {
  if constexpr (is_h16<T>::value) return T::fmt::to_f(v.bits);
  else return (float)v;
}
template <typename T, typename R>
__device__ __forceinline__ R load(const T* p) {
  if constexpr (is_h16<T>::value) return (R)to_f(*p);
  else return (R)*p;
}
template <typename T, typename R>
__device__ __forceinline__ T store(R v) {
  if constexpr (is_h16<T>::value) return T{T::fmt::from_f((float)v)};  // nearest even
  else return (T)v;
}

template <typename R>
__device__ __forceinline__ R tanh_r(R x) {
  if constexpr (std::is_same<R, double>::value) {
    return tanh(x);
  } else {
    // branch-free: 1 - 2 / (2^(2 x log2 e) + 1) with v_exp_f32 / v_rcp_f32
    const float e = __builtin_amdgcn_exp2f(x * 2.8853900817779268f);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
  }
}

constexpr int kMaxD3 = 256, kMaxM2 = 16;
constexpr int kWaves = 2;  // wavefronts (rows) per workgroup of the descriptor kernels

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// row copy global <-> LDS in 16-byte pieces when the row allows it
__device__ __forceinline__ void copy_row(void* dst, const void* src, int bytes, int lane) {
  if ((bytes & 15) == 0 && ((reinterpret_cast<size_t>(src) | reinterpret_cast<size_t>(dst)) & 15) == 0) {
    const int4* s4 = static_cast<const int4*>(src);
    int4* d4 = static_cast<int4*>(dst);
    for (int q = lane; q < bytes / 16; q += 64) d4[q] = s4[q];
  } else {
    const unsigned short* s2 = static_cast<const unsigned short*>(src);
    unsigned short* d2 = static_cast<unsigned short*>(dst);
    for (int q = lane; q < bytes / 2; q += 64) d2[q] = s2[q];
  }
}

// Per-wavefront LDS: xyz [4*d3] (R), then a descriptor row [d3*m2] (T).
template <typename R, typename T>
__host__ __device__ constexpr size_t desc_lds_per_wave(int d3, int m2) {
  return ((sizeof(R) * 4 * d3 + 15) / 16) * 16 + ((sizeof(T) * d3 * m2 + 15) / 16) * 16 +
         sizeof(R) * 4 * d3;
}

// D rows in slot order; one wavefront per slot.  The row is formed in LDS and
// written with 16-byte stores.
template <typename R, typename T>
__global__ void __launch_bounds__(64 * kWaves) k_desc_fwd(T* __restrict__ desc,
                                                          const R* __restrict__ xyz,
                                                          const int* __restrict__ order,
                                                          int nslots, int d3, int m2) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = blockIdx.x * kWaves + wave;
  if (slot >= nslots) return;  // no workgroup barrier below
  unsigned char* base = lds + wave * desc_lds_per_wave<R, T>(d3, m2);
  R* xw = reinterpret_cast<R*>(base);
  T* dw = reinterpret_cast<T*>(base + ((sizeof(R) * 4 * d3 + 15) / 16) * 16);
  copy_row(xw, xyz + (size_t)order[slot] * 4 * d3, (int)sizeof(R) * 4 * d3, lane);
  wave_sync_lds();
  for (int q = lane; q < d3 * m2; q += 64) {
    const int k = q / m2, m = q - k * m2;
    const R v = xw[k] * xw[m] + xw[d3 + k] * xw[d3 + m] + xw[2 * d3 + k] * xw[2 * d3 + m] +
                xw[3 * d3 + k] * xw[3 * d3 + m];
    dw[q] = store<T, R>(v);
  }
  wave_sync_lds();
  copy_row(desc + (size_t)slot * d3 * m2, dw, (int)sizeof(T) * d3 * m2, lane);
}

template <typename R, typename T>
__global__ void __launch_bounds__(64 * kWaves) k_desc_bwd(R* __restrict__ gxyz,
                                                          const T* __restrict__ gd,
                                                          const R* __restrict__ xyz,
                                                          const int* __restrict__ order,
                                                          int nslots, int d3, int m2) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int slot = blockIdx.x * kWaves + wave;
  if (slot >= nslots) return;
  const int atom = order[slot];
  unsigned char* base = lds + wave * desc_lds_per_wave<R, T>(d3, m2);
  R* xw = reinterpret_cast<R*>(base);
  T* g = reinterpret_cast<T*>(base + ((sizeof(R) * 4 * d3 + 15) / 16) * 16);
  R* gw = reinterpret_cast<R*>(base + ((sizeof(R) * 4 * d3 + 15) / 16) * 16 +
                               ((sizeof(T) * d3 * m2 + 15) / 16) * 16);
  copy_row(xw, xyz + (size_t)atom * 4 * d3, (int)sizeof(R) * 4 * d3, lane);
  copy_row(g, gd + (size_t)slot * d3 * m2, (int)sizeof(T) * d3 * m2, lane);
  wave_sync_lds();
  // first term: lanes over p
  for (int p = lane; p < d3; p += 64) {
    R a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (int m = 0; m < m2; ++m) {
      const R v = load<T, R>(g + p * m2 + m);
      a0 += v * xw[m];
      a1 += v * xw[d3 + m];
      a2 += v * xw[2 * d3 + m];
      a3 += v * xw[3 * d3 + m];
    }
    gw[p] = a0;
    gw[d3 + p] = a1;
    gw[2 * d3 + p] = a2;
    gw[3 * d3 + p] = a3;
  }
  wave_sync_lds();
  // second term: lane (c, p), p < m2
  {
    const int cc = lane / kMaxM2, p = lane % kMaxM2;
    if (p < m2) {
      R b = 0;
      for (int k = 0; k < d3; ++k) b += load<T, R>(g + k * m2 + p) * xw[cc * d3 + k];
      gw[cc * d3 + p] += b;
    }
  }
  wave_sync_lds();
  copy_row(gxyz + (size_t)atom * 4 * d3, gw, (int)sizeof(R) * 4 * d3, lane);
}

// y = act(z + b) * idt + res(y_prev); yg = y in the GEMM type; sd = idt act'(z + b)
// (tanh: idt (1 - t^2); other activations through act_fd when GEN).
// One wavefront per row (no index division), lanes over channels.
template <typename R, typename T, bool GEN>
__global__ void __launch_bounds__(256) k_layer_fwd(R* __restrict__ y, T* __restrict__ yg,
                                                   T* __restrict__ sd, const T* __restrict__ z,
                                                   const R* __restrict__ bias,
                                                   const R* __restrict__ idt,
                                                   const R* __restrict__ y_prev, int n, int h,
                                                   int hin, int res, ActSpec act) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const size_t o = (size_t)row * h;
  const R* yp = y_prev ? y_prev + (size_t)row * hin : nullptr;
  for (int ch = lane; ch < h; ch += 64) {
    const R zb = load<T, R>(z + o + ch) + bias[ch];
    R t, dt;
    if constexpr (GEN) {
      act_fd<R>(act, zb, t, dt);
    } else {
      t = tanh_r<R>(zb);
      dt = R(1) - t * t;
    }
    const R s = idt ? idt[ch] : R(1);
    R v = t * s;
    if (res == 1) v += yp[ch];
    else if (res == 2) v += yp[ch < hin ? ch : ch - hin];
    y[o + ch] = v;
    if (yg) yg[o + ch] = store<T, R>(v);
    sd[o + ch] = store<T, R>(s * dt);
  }
}

// atom_e[order[i]] = y_i . w_out + (b_out + bias_atom_e[type]); gz = w_out * sd
template <typename R, typename T>
__global__ void __launch_bounds__(256) k_energy(double* __restrict__ atom_e, T* __restrict__ gz,
                                                const R* __restrict__ y, const T* __restrict__ sd,
                                                const R* __restrict__ wout, double ebias,
                                                const int* __restrict__ order, int n, int h) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= n) return;
  R acc = 0;
  for (int ch = lane; ch < h; ch += 64) {
    const size_t q = (size_t)row * h + ch;
    acc += y[q] * wout[ch];
    gz[q] = store<T, R>(wout[ch] * load<T, R>(sd + q));
  }
  acc = wave_sum(acc);
  if (lane == 0) atom_e[order[row]] = (double)acc + ebias;
}

// gy = mm + res_next^T(gy_next | w_out); gz = gy * sd.  One wavefront per row.
template <typename R, typename T>
__global__ void __launch_bounds__(256) k_layer_bwd(R* __restrict__ gy, T* __restrict__ gz,
                                                   const T* __restrict__ mm,
                                                   const R* __restrict__ gy_next,
                                                   const R* __restrict__ wvec,
                                                   const T* __restrict__ sd, int n, int h,
                                                   int hnext, int res_next) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= n) return;
  const size_t o = (size_t)row * h;
  const R* gn = wvec ? wvec : (gy_next ? gy_next + (size_t)row * hnext : nullptr);
  for (int ch = lane; ch < h; ch += 64) {
    R v = load<T, R>(mm + o + ch);
    if (res_next != 0) {
      // gradient of the next layer's residual copy of this layer's output
      v += gn[ch];
      if (res_next == 2) v += gn[ch + h];
    }
    if (gy) gy[o + ch] = v;
    if (gz) gz[o + ch] = store<T, R>(v * load<T, R>(sd + o + ch));
  }
}

// ---------------------------------------------------------------------------
// Vectorised forms (widths divisible by 4): each thread moves 4 consecutive
// channels with one 8/16/32-byte access per array.
template <typename R>
struct V4 {
  R v[4];
};
template <typename T, typename R>
__device__ __forceinline__ V4<R> load4(const T* p) {
  V4<R> r;
  if constexpr (is_h16<T>::value) {
    using E = typename T::fmt;
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    r.v[0] = (R)E::lo(u.x);
    r.v[1] = (R)E::hi(u.x);
    r.v[2] = (R)E::lo(u.y);
    r.v[3] = (R)E::hi(u.y);
  } else if constexpr (std::is_same<T, float>::value) {
    const float4 f = *reinterpret_cast<const float4*>(p);
    r.v[0] = f.x; r.v[1] = f.y; r.v[2] = f.z; r.v[3] = f.w;
  } else {
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    r.v[0] = a.x; r.v[1] = a.y; r.v[2] = b.x; r.v[3] = b.y;
  }
  return r;
}
template <typename T, typename R>
__device__ __forceinline__ void store4(T* p, const V4<R>& r) {
  if constexpr (is_h16<T>::value) {
    using E = typename T::fmt;
    uint2 u;
    u.x = E::pack2((float)r.v[0], (float)r.v[1]);
    u.y = E::pack2((float)r.v[2], (float)r.v[3]);
    *reinterpret_cast<uint2*>(p) = u;
  } else if constexpr (std::is_same<T, float>::value) {
    *reinterpret_cast<float4*>(p) = make_float4(r.v[0], r.v[1], r.v[2], r.v[3]);
  } else {
    reinterpret_cast<double2*>(p)[0] = make_double2(r.v[0], r.v[1]);
    reinterpret_cast<double2*>(p)[1] = make_double2(r.v[2], r.v[3]);
  }
}

// k_layer_fwd, one thread per 4 channels of a row (h, hin divisible by 4)
template <typename R, typename T, bool GEN>
__global__ void __launch_bounds__(256) k_layer_fwd4(R* __restrict__ y, T* __restrict__ yg,
                                                    T* __restrict__ sd, const T* __restrict__ z,
                                                    const R* __restrict__ bias,
                                                    const R* __restrict__ idt,
                                                    const R* __restrict__ y_prev, int n, int h,
                                                    int hin, int res, ActSpec act) {
  const int h4 = h >> 2;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)n * h4) return;
  const int row = (int)(q / h4), ch = 4 * (int)(q - (long long)row * h4);
  const size_t o = (size_t)row * h + ch;
  const V4<R> zz = load4<T, R>(z + o);
  const V4<R> bb = load4<R, R>(bias + ch);
  V4<R> ss;
  if (idt) ss = load4<R, R>(idt + ch);
  else ss.v[0] = ss.v[1] = ss.v[2] = ss.v[3] = R(1);
  V4<R> pv;
  if (res != 0) pv = load4<R, R>(y_prev + (size_t)row * hin + (res == 2 && ch >= hin ? ch - hin : ch));
  V4<R> yv, dv;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    R t, dt;
    if constexpr (GEN) {
      act_fd<R>(act, zz.v[k] + bb.v[k], t, dt);
    } else {
      t = tanh_r<R>(zz.v[k] + bb.v[k]);
      dt = R(1) - t * t;
    }
    yv.v[k] = t * ss.v[k] + (res != 0 ? pv.v[k] : R(0));
    dv.v[k] = ss.v[k] * dt;
  }
  store4<R, R>(y + o, yv);
  if (yg) store4<T, R>(yg + o, yv);
  store4<T, R>(sd + o, dv);
}

// k_energy with 4 channels per lane access
template <typename R, typename T>
__global__ void __launch_bounds__(256) k_energy4(double* __restrict__ atom_e, T* __restrict__ gz,
                                                 const R* __restrict__ y,
                                                 const T* __restrict__ sd,
                                                 const R* __restrict__ wout, double ebias,
                                                 const int* __restrict__ order, int n, int h) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + wave;
  if (row >= n) return;
  R acc = 0;
  for (int ch = 4 * lane; ch < h; ch += 256) {
    const size_t o = (size_t)row * h + ch;
    const V4<R> yv = load4<R, R>(y + o), wv = load4<R, R>(wout + ch), sv = load4<T, R>(sd + o);
    V4<R> g;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc += yv.v[k] * wv.v[k];
      g.v[k] = wv.v[k] * sv.v[k];
    }
    store4<T, R>(gz + o, g);
  }
  acc = wave_sum(acc);
  if (lane == 0) atom_e[order[row]] = (double)acc + ebias;
}

// k_layer_bwd, one thread per 4 channels of a row (h divisible by 4)
template <typename R, typename T>
__global__ void __launch_bounds__(256) k_layer_bwd4(R* __restrict__ gy, T* __restrict__ gz,
                                                    const T* __restrict__ mm,
                                                    const R* __restrict__ gy_next,
                                                    const R* __restrict__ wvec,
                                                    const T* __restrict__ sd, int n, int h,
                                                    int hnext, int res_next) {
  const int h4 = h >> 2;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)n * h4) return;
  const int row = (int)(q / h4), ch = 4 * (int)(q - (long long)row * h4);
  const size_t o = (size_t)row * h + ch;
  V4<R> v = load4<T, R>(mm + o);
  if (res_next != 0) {
    const R* gn = wvec ? wvec : gy_next + (size_t)row * hnext;
    const V4<R> a = load4<R, R>(gn + ch);
#pragma unroll
    for (int k = 0; k < 4; ++k) v.v[k] += a.v[k];
    if (res_next == 2) {
      const V4<R> b = load4<R, R>(gn + ch + h);
#pragma unroll
      for (int k = 0; k < 4; ++k) v.v[k] += b.v[k];
    }
  }
  if (gy) store4<R, R>(gy + o, v);
  if (gz) {
    const V4<R> sv = load4<T, R>(sd + o);
#pragma unroll
    for (int k = 0; k < 4; ++k) v.v[k] *= sv.v[k];
    store4<T, R>(gz + o, v);
  }
}

// 16 values of a descriptor-gradient row (16-bit or FP32) as float.
template <typename T>
__device__ __forceinline__ void load_row16(const T* __restrict__ p, float v[16]) {
  if constexpr (is_h16<T>::value) {
    using E = typename T::fmt;
    const uint4 a = reinterpret_cast<const uint4*>(p)[0], b = reinterpret_cast<const uint4*>(p)[1];
    const unsigned u[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[2 * k] = E::lo(u[k]);
      v[2 * k + 1] = E::hi(u[k]);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float4 f = reinterpret_cast<const float4*>(p)[k];
      v[4 * k] = f.x; v[4 * k + 1] = f.y; v[4 * k + 2] = f.z; v[4 * k + 3] = f.w;
    }
  }
}

// k_desc_bwd for M2 = 16 without LDS: one wavefront per slot, lane l takes
// rows p = l + 64 j (d3 <= 64 NP).  The first term is the lane's own rows;
// the second, sum_k dE/dD[k][m] xyz[c][k] for (c, m) in 4 x 16, is summed
// over the lanes by recursive halving (lane l ends with entry l), then
// gathered by the lanes p < 16.
template <typename T, int NP>
__global__ void __launch_bounds__(256) k_desc_bwd_r(float* __restrict__ gxyz,
                                                    const T* __restrict__ gd,
                                                    const float* __restrict__ xyz,
                                                    const int* __restrict__ order, int nslots,
                                                    int d3) {
  constexpr int M2 = 16;
  // wave-uniform indices (so xyz[c][m < 16] are scalar loads)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int slot = blockIdx.x * 4 + wave;
  if (slot >= nslots) return;
  const int atom = __builtin_amdgcn_readfirstlane(order[slot]);
  const float* x = xyz + (size_t)atom * 4 * d3;
  const T* g = gd + (size_t)slot * d3 * M2;
  float xm[4][M2];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int m = 0; m < M2; ++m) xm[c][m] = x[c * d3 + m];
  float part[4 * M2];
#pragma unroll
  for (int k = 0; k < 4 * M2; ++k) part[k] = 0.f;
  float t1[NP][4];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = lane + 64 * j;
    const bool ok = p < d3;
    float gv[M2], xp[4];
    load_row16<T>(g + (size_t)(ok ? p : 0) * M2, gv);
#pragma unroll
    for (int m = 0; m < M2; ++m) gv[m] = ok ? gv[m] : 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) xp[c] = ok ? x[c * d3 + p] : 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      float a = 0.f;
#pragma unroll
      for (int m = 0; m < M2; ++m) a = __builtin_fmaf(gv[m], xm[c][m], a);
      t1[j][c] = a;
#pragma unroll
      for (int m = 0; m < M2; ++m) part[c * M2 + m] = __builtin_fmaf(gv[m], xp[c], part[c * M2 + m]);
    }
  }
#pragma unroll
  for (int sh = 32; sh >= 1; sh >>= 1) {
    const bool upper = (lane & sh) != 0;
#pragma unroll
    for (int i = 0; i < sh; ++i) {
      const float send = upper ? part[i] : part[i + sh];
      const float keep = upper ? part[i + sh] : part[i];
      part[i] = keep + __shfl_xor(send, sh, 64);
    }
  }
  float add[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) add[c] = __shfl(part[0], c * M2 + (lane & (M2 - 1)), 64);
  float* o = gxyz + (size_t)atom * 4 * d3;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = lane + 64 * j;
    if (p < d3)
#pragma unroll
      for (int c = 0; c < 4; ++c) o[c * d3 + p] = t1[j][c] + ((j == 0 && lane < M2) ? add[c] : 0.f);
  }
}

// Descriptor rows without LDS staging: one wavefront per slot, lanes over
// channels p (stride 64).  xyz[c][m] for m < M2 is wave-uniform (scalar
// loads); the lane's own xyz[c][p] and its D row piece D[p][0..M2) stay in
// registers and the row piece is written with vector stores.
// four values rounded to 16 bits (nearest even) in two words
template <typename T, typename R>
__device__ __forceinline__ uint2 pack4_h16(const V4<R>& r) {
  using E = typename T::fmt;
  return make_uint2(E::pack2((float)r.v[0], (float)r.v[1]), E::pack2((float)r.v[2], (float)r.v[3]));
}

template <typename R, typename T, int M2>
__global__ void __launch_bounds__(256) k_desc_fwd_r(T* __restrict__ desc,
                                                    const R* __restrict__ xyz,
                                                    const int* __restrict__ order, int nslots,
                                                    int d3) {
  // wave-uniform indices (so xyz[c][m < M2] are scalar loads)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int slot = blockIdx.x * 4 + wave;
  if (slot >= nslots) return;
  const R* x = xyz + (size_t)__builtin_amdgcn_readfirstlane(order[slot]) * 4 * d3;
  R xm[4][M2];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int m = 0; m < M2; ++m) xm[c][m] = x[c * d3 + m];
  T* drow = desc + (size_t)slot * d3 * M2;
  for (int p = lane; p < d3; p += 64) {
    const R x0 = x[p], x1 = x[d3 + p], x2 = x[2 * d3 + p], x3 = x[3 * d3 + p];
    V4<R> pend;
#pragma unroll
    for (int m0 = 0; m0 < M2; m0 += 4) {
      V4<R> o;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int m = m0 + k;
        o.v[k] = x0 * xm[0][m] + x1 * xm[1][m] + x2 * xm[2][m] + x3 * xm[3][m];
      }
      if constexpr (std::is_same<T, f16x2s>::value) {
        // hi and lo parts, each a pair of 4-value groups per 16-byte store
        static_assert(M2 % 8 == 0, "split rows need M2 % 8 == 0");
        if ((m0 & 4) == 0) {
          pend = o;
        } else {
          using E = h16::F16;
          const V4<R> a = pend, b = o;
          const uint2 h0 = pack4_h16<f16s>(a), h1 = pack4_h16<f16s>(b);
          V4<R> la, lb;
          la.v[0] = a.v[0] - E::lo(h0.x); la.v[1] = a.v[1] - E::hi(h0.x);
          la.v[2] = a.v[2] - E::lo(h0.y); la.v[3] = a.v[3] - E::hi(h0.y);
          lb.v[0] = b.v[0] - E::lo(h1.x); lb.v[1] = b.v[1] - E::hi(h1.x);
          lb.v[2] = b.v[2] - E::lo(h1.y); lb.v[3] = b.v[3] - E::hi(h1.y);
          const uint2 l0 = pack4_h16<f16s>(la), l1 = pack4_h16<f16s>(lb);
          T* hp = desc + (size_t)slot * 2 * d3 * M2 + p * M2 + m0 - 4;
          *reinterpret_cast<uint4*>(hp) = make_uint4(h0.x, h0.y, h1.x, h1.y);
          *reinterpret_cast<uint4*>(hp + d3 * M2) = make_uint4(l0.x, l0.y, l1.x, l1.y);
        }
      } else if constexpr (is_h16<T>::value && M2 % 8 == 0) {
        // 16-bit: pairs of 4-value groups as one 16-byte store
        if ((m0 & 4) == 0) {
          pend = o;
        } else {
          const uint2 lo = pack4_h16<T>(pend), hi = pack4_h16<T>(o);
          *reinterpret_cast<uint4*>(drow + p * M2 + m0 - 4) = make_uint4(lo.x, lo.y, hi.x, hi.y);
        }
      } else {
        store4<T, R>(drow + p * M2 + m0, o);
      }
    }
  }
}

}  // namespace fit
}  // namespace
}  // namespace dpk

using namespace dpk;

#define DPK_FIT_DISPATCH(prec, CALL)                                                   \
  switch (prec) {                                                                      \
    case DPK_PREC_F64: {                                                               \
      using R = double;                                                                \
      using T = double;                                                                \
      CALL;                                                                            \
      break;                                                                           \
    }                                                                                  \
    case DPK_PREC_F32: {                                                               \
      using R = float;                                                                 \
      using T = float;                                                                 \
      CALL;                                                                            \
      break;                                                                           \
    }                                                                                  \
    case DPK_PREC_BF16: {                                                              \
      using R = float;                                                                 \
      using T = fit::bf16s;                                                            \
      CALL;                                                                            \
      break;                                                                           \
    }                                                                                  \
    case DPK_PREC_F16: {                                                               \
      using R = float;                                                                 \
      using T = fit::f16s;                                                             \
      CALL;                                                                            \
      break;                                                                           \
    }                                                                                  \
    default:                                                                           \
      set_error("fitting: unknown precision %d", prec);                               \
      return DPK_ERR_ARG;                                                              \
  }

extern "C" int dpk_fit_descriptor_fwd(void* desc, int prec, const void* xyz, const int* order,
                                      int nslots, int d3, int m2, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nslots >= 0 && d3 > 0 && d3 <= fit::kMaxD3 && m2 > 0 && m2 <= fit::kMaxM2 &&
                  m2 <= d3,
              DPK_ERR_UNSUPPORTED, "fit_descriptor_fwd: d3=%d m2=%d (d3<=%d, m2<=%d)", d3, m2,
              fit::kMaxD3, fit::kMaxM2);
  if (nslots == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  const size_t lds_f64 = fit::kWaves * fit::desc_lds_per_wave<double, double>(d3, m2);
  DPK_REQUIRE(prec != DPK_PREC_F64 || lds_f64 <= 65536, DPK_ERR_UNSUPPORTED,
              "fit_descriptor_fwd: d3=%d m2=%d needs %zu B of LDS in FP64", d3, m2, lds_f64);
  if (prec == DPK_PREC_F16X2) {
    DPK_REQUIRE(m2 == 16 || m2 == 8, DPK_ERR_UNSUPPORTED,
                "fit_descriptor_fwd: split FP16 rows need m2 = 8 or 16 (got %d)", m2);
    if (m2 == 16)
      hipLaunchKernelGGL((fit::k_desc_fwd_r<float, fit::f16x2s, 16>), dim3(cdiv(nslots, 4)),
                         dim3(256), 0, st, static_cast<fit::f16x2s*>(desc),
                         static_cast<const float*>(xyz), order, nslots, d3);
    else
      hipLaunchKernelGGL((fit::k_desc_fwd_r<float, fit::f16x2s, 8>), dim3(cdiv(nslots, 4)),
                         dim3(256), 0, st, static_cast<fit::f16x2s*>(desc),
                         static_cast<const float*>(xyz), order, nslots, d3);
    DPK_LAUNCHED();
    return DPK_OK;
  }
  if (m2 == 16 || m2 == 8 || m2 == 4) {
#define DPK_DESC_FWD(M)                                                                     \
  DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_desc_fwd_r<R, T, M>), dim3(cdiv(nslots, 4)), \
                                            dim3(256), 0, st, static_cast<T*>(desc),          \
                                            static_cast<const R*>(xyz), order, nslots, d3))
    if (m2 == 16) {
      DPK_DESC_FWD(16);
    } else if (m2 == 8) {
      DPK_DESC_FWD(8);
    } else {
      DPK_DESC_FWD(4);
    }
#undef DPK_DESC_FWD
    DPK_LAUNCHED();
    return DPK_OK;
  }
  DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_desc_fwd<R, T>),
                                            dim3(cdiv(nslots, fit::kWaves)), dim3(64 * fit::kWaves),
                                            (fit::kWaves * fit::desc_lds_per_wave<R, T>(d3, m2)), st,
                                            static_cast<T*>(desc),
                                            static_cast<const R*>(xyz), order, nslots, d3, m2));
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_fit_descriptor_bwd(void* gxyz, int prec, const void* gdesc, const void* xyz,
                                      const int* order, int nslots, int d3, int m2,
                                      dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nslots >= 0 && d3 > 0 && d3 <= fit::kMaxD3 && m2 > 0 && m2 <= fit::kMaxM2 &&
                  m2 <= d3,
              DPK_ERR_UNSUPPORTED, "fit_descriptor_bwd: d3=%d m2=%d (d3<=%d, m2<=%d)", d3, m2,
              fit::kMaxD3, fit::kMaxM2);
  if (nslots == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  const size_t lds_f64 = fit::kWaves * fit::desc_lds_per_wave<double, double>(d3, m2);
  DPK_REQUIRE(prec != DPK_PREC_F64 || lds_f64 <= 65536, DPK_ERR_UNSUPPORTED,
              "fit_descriptor_bwd: d3=%d m2=%d needs %zu B of LDS in FP64", d3, m2, lds_f64);
  if (prec == DPK_PREC_F16X2) prec = DPK_PREC_F32;  // split mode: dE/dD is FP32
  if (m2 == 16 && prec != DPK_PREC_F64) {
    const int np = (d3 + 63) / 64;
#define DPK_DESC_BWD_R(T, NP)                                                                    \
  hipLaunchKernelGGL((fit::k_desc_bwd_r<T, NP>), dim3(cdiv(nslots, 4)), dim3(256), 0, st,         \
                     static_cast<float*>(gxyz), static_cast<const T*>(gdesc),                    \
                     static_cast<const float*>(xyz), order, nslots, d3)
#define DPK_DESC_BWD_NP(T)      \
  if (np == 1) DPK_DESC_BWD_R(T, 1); \
  else if (np == 2) DPK_DESC_BWD_R(T, 2); \
  else if (np == 3) DPK_DESC_BWD_R(T, 3); \
  else DPK_DESC_BWD_R(T, 4);
    if (prec == DPK_PREC_BF16) {
      DPK_DESC_BWD_NP(fit::bf16s)
    } else if (prec == DPK_PREC_F16) {
      DPK_DESC_BWD_NP(fit::f16s)
    } else {
      DPK_DESC_BWD_NP(float)
    }
#undef DPK_DESC_BWD_NP
#undef DPK_DESC_BWD_R
    DPK_LAUNCHED();
    return DPK_OK;
  }
  DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_desc_bwd<R, T>),
                                            dim3(cdiv(nslots, fit::kWaves)), dim3(64 * fit::kWaves),
                                            (fit::kWaves * fit::desc_lds_per_wave<R, T>(d3, m2)), st,
                                            static_cast<R*>(gxyz),
                                            static_cast<const T*>(gdesc),
                                            static_cast<const R*>(xyz), order, nslots, d3, m2));
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_fit_layer_fwd(void* y, void* yg, void* sd, int prec, const void* z,
                                 const void* bias, const void* idt, const void* y_prev, int n,
                                 int h, int hin, int res, int act, double act_param,
                                 dpk_stream_t stream) {
  clear_error();
  ActSpec as;
  DPK_REQUIRE(act_spec(act, act_param, &as), DPK_ERR_ARG, "fit_layer_fwd: unknown activation %d",
              act);
  DPK_REQUIRE(n >= 0 && h > 0 && res >= 0 && res <= 2, DPK_ERR_ARG, "fit_layer_fwd: bad args");
  DPK_REQUIRE(res == 0 || (y_prev && (res == 1 ? hin == h : 2 * hin == h)), DPK_ERR_ARG,
              "fit_layer_fwd: resnet %d needs y_prev with width %d (h=%d)", res, hin, h);
  DPK_REQUIRE(y && sd && z && bias, DPK_ERR_ARG, "fit_layer_fwd: null buffer");
  if (n == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  if (h % 4 == 0 && (res != 2 || hin % 4 == 0)) {
    const long long quads = (long long)n * (h / 4);
    if (act == DPK_ACT_TANH) {
      DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_layer_fwd4<R, T, false>),
                                                dim3(cdiv(quads, 256)), dim3(256), 0, st,
                                                static_cast<R*>(y), static_cast<T*>(yg),
                                                static_cast<T*>(sd), static_cast<const T*>(z),
                                                static_cast<const R*>(bias),
                                                static_cast<const R*>(idt),
                                                static_cast<const R*>(y_prev), n, h, hin, res, as));
    } else {
      DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_layer_fwd4<R, T, true>),
                                                dim3(cdiv(quads, 256)), dim3(256), 0, st,
                                                static_cast<R*>(y), static_cast<T*>(yg),
                                                static_cast<T*>(sd), static_cast<const T*>(z),
                                                static_cast<const R*>(bias),
                                                static_cast<const R*>(idt),
                                                static_cast<const R*>(y_prev), n, h, hin, res, as));
    }
    DPK_LAUNCHED();
    return DPK_OK;
  }
  if (act == DPK_ACT_TANH) {
    DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_layer_fwd<R, T, false>), dim3(cdiv(n, 4)),
                                              dim3(256), 0, st, static_cast<R*>(y),
                                              static_cast<T*>(yg), static_cast<T*>(sd),
                                              static_cast<const T*>(z), static_cast<const R*>(bias),
                                              static_cast<const R*>(idt),
                                              static_cast<const R*>(y_prev), n, h, hin, res, as));
  } else {
    DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_layer_fwd<R, T, true>), dim3(cdiv(n, 4)),
                                              dim3(256), 0, st, static_cast<R*>(y),
                                              static_cast<T*>(yg), static_cast<T*>(sd),
                                              static_cast<const T*>(z), static_cast<const R*>(bias),
                                              static_cast<const R*>(idt),
                                              static_cast<const R*>(y_prev), n, h, hin, res, as));
  }
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_fit_energy(double* atom_energy, void* gz, int prec, const void* y,
                              const void* sd, const void* wout, double ebias, const int* order,
                              int n, int h, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(n >= 0 && h > 0, DPK_ERR_ARG, "fit_energy: bad sizes");
  DPK_REQUIRE(atom_energy && gz && y && sd && wout && order, DPK_ERR_ARG, "fit_energy: null buffer");
  if (n == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  if (h % 4 == 0) {
    DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_energy4<R, T>), dim3(cdiv(n, 4)), dim3(256),
                                              0, st, atom_energy, static_cast<T*>(gz),
                                              static_cast<const R*>(y), static_cast<const T*>(sd),
                                              static_cast<const R*>(wout), ebias, order, n, h));
    DPK_LAUNCHED();
    return DPK_OK;
  }
  DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_energy<R, T>), dim3(cdiv(n, 4)), dim3(256), 0,
                                            st, atom_energy, static_cast<T*>(gz),
                                            static_cast<const R*>(y), static_cast<const T*>(sd),
                                            static_cast<const R*>(wout), ebias, order, n, h));
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_fit_layer_bwd(void* gy, void* gz, int prec, const void* mm,
                                 const void* gy_next, const void* wvec, const void* sd, int n,
                                 int h, int hnext, int res_next, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(n >= 0 && h > 0 && res_next >= 0 && res_next <= 2, DPK_ERR_ARG,
              "fit_layer_bwd: bad args");
  DPK_REQUIRE(res_next == 0 || ((gy_next || wvec) && (res_next == 1 ? hnext == h : hnext == 2 * h)),
              DPK_ERR_ARG, "fit_layer_bwd: resnet %d needs the next gradient (hnext=%d, h=%d)",
              res_next, hnext, h);
  DPK_REQUIRE(mm && (gz == nullptr || sd), DPK_ERR_ARG, "fit_layer_bwd: null buffer");
  if (n == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  if (h % 4 == 0) {
    const long long quads = (long long)n * (h / 4);
    DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_layer_bwd4<R, T>), dim3(cdiv(quads, 256)),
                                              dim3(256), 0, st, static_cast<R*>(gy),
                                              static_cast<T*>(gz), static_cast<const T*>(mm),
                                              static_cast<const R*>(gy_next),
                                              static_cast<const R*>(wvec),
                                              static_cast<const T*>(sd), n, h, hnext, res_next));
    DPK_LAUNCHED();
    return DPK_OK;
  }
  DPK_FIT_DISPATCH(prec, hipLaunchKernelGGL((fit::k_layer_bwd<R, T>), dim3(cdiv(n, 4)), dim3(256),
                                            0, st, static_cast<R*>(gy), static_cast<T*>(gz),
                                            static_cast<const T*>(mm),
                                            static_cast<const R*>(gy_next),
                                            static_cast<const R*>(wvec), static_cast<const T*>(sd),
                                            n, h, hnext, res_next));
  DPK_LAUNCHED();
  return DPK_OK;
}
