// Shared helpers for the libdpk HIP sources (gfx950, wave64).
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "dpk.h"

namespace dpk {

// ------------------------------------------------------------ error state --
void set_error(const char* fmt, ...);
void clear_error();
int tune(int key);  // current value of a DPK_TUNE_* tunable

#define DPK_REQUIRE(cond, code, ...) \
  do {                               \
    if (!(cond)) {                   \
      ::dpk::set_error(__VA_ARGS__); \
      return (code);                 \
    }                                \
  } while (0)

#define DPK_HIP(call)                                                      \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      ::dpk::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                \
      return DPK_ERR_HIP;                                                  \
    }                                                                      \
  } while (0)

#define DPK_LAUNCHED()                                                      \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess) {                                                 \
      ::dpk::set_error("kernel launch failed: %s (%s:%d)",                  \
                       hipGetErrorString(e_), __FILE__, __LINE__);          \
      return DPK_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)

inline hipStream_t as_stream(dpk_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned int cdiv(long long a, long long b) { return (unsigned int)((a + b - 1) / b); }

constexpr int kWave = 64;

// ------------------------------------------------------------ device math --

// |d| with the rounding torch's CPU vector norm uses for a 3-vector:
// sqrt(fma(z, z, fma(y, y, x*x))), IEEE sqrt.
__device__ __forceinline__ double norm3_ref(double x, double y, double z) {
  double acc = __dmul_rn(x, x);
  acc = __fma_rn(y, y, acc);
  acc = __fma_rn(z, z, acc);
  return __dsqrt_rn(acc);
}

// Division by a run-time constant d (1 <= d <= 2^31) as a multiply-high and a
// shift (Granlund & Montgomery, "Division by invariant integers using
// multiplication", 1994: round-up variant, m = floor(2^32 (2^s - d) / d) + 1
// with s = ceil(log2 d)); exact for 0 <= n < 2^31.
struct FastDiv {
  unsigned d, m, s;
  __host__ __device__ void init(unsigned dd) {
    d = dd;
    s = 0;
    while ((1ull << s) < dd) ++s;
    m = (unsigned)((((1ull << 32) * ((1ull << s) - dd)) / dd) + 1);
  }
  __device__ __forceinline__ unsigned div(unsigned n) const {
    return (__umulhi(n, m) + n) >> s;
  }
};

// 1/x in FP64 from the hardware estimate and two Newton steps (for operands
// whose result is rounded to FP32 afterwards; within 1 ulp of 1.0 / x).
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = __fma_rn(y, __fma_rn(-x, y, 1.0), y);
  y = __fma_rn(y, __fma_rn(-x, y, 1.0), y);
  return y;
}

// Smooth switch of deepmd/pt/utils/preprocess.py:9-18 (value) and
// source/lib/include/switcher.h:60-83 (value + derivative).
template <typename T>
__device__ __forceinline__ void spline5(T r, T rmin, T rmax, T& sw, T& dsw) {
  if (r < rmin) {
    sw = T(1);
    dsw = T(0);
  } else if (r < rmax) {
    T du = T(1) / (rmax - rmin);
    T u = (r - rmin) * du;
    T u2 = u * u;
    sw = u2 * u * (T(-6) * u2 + T(15) * u - T(10)) + T(1);
    dsw = (T(3) * u2 * (T(-6) * u2 + T(15) * u - T(10)) + u2 * u * (T(-12) * u + T(15))) * du;
  } else {
    sw = T(0);
    dsw = T(0);
  }
}

// Env-matrix row of a neighbour at displacement d = (x, y, z) = r_j - r_i,
// |d| = r, with u = r + protection (deepmd/pt/model/descriptor/env_mat.py:
// 38-45): R = sw(r) (1/u, x/u^2, y/u^2, z/u^2).  Given iu = 1/u and
// inr = 1/r, env_centre_deriv writes dR_c/d r_i[a] (a = x, y, z) contracted
// with g (4 values): f[a] = sum_c g[c] dR_c/dr_i[a] (env_mat.cc:176-207 at
// protection 0, where iu = inr).
template <typename T>
__device__ __forceinline__ void env_centre_deriv(const T d[3], T inr, T iu, T sw, T dsw,
                                                 const T g[4], T f[3]) {
  const T iu2 = iu * iu;
  const T c0 = inr * iu * (iu * sw - dsw);  // dR0/dr_i = c0 d
  const T c1 = T(2) * iu2 * iu * inr;
  const T v[3] = {d[0] * iu2, d[1] * iu2, d[2] * iu2};
  for (int a = 0; a < 3; ++a) {
    T acc = g[0] * (d[a] * c0);
    for (int c = 0; c < 3; ++c) {
      const T t = c1 * d[c] * d[a] - (c == a ? iu2 : T(0));
      acc += g[c + 1] * (t * sw - v[c] * dsw * d[a] * inr);
    }
    f[a] = acc;
  }
}

// Activation functions of deepmd/pt/utils/utils.py:174-212 (ActivationFn),
// value f(z) and derivative f'(z), with torch's conventions at the kinks
// (relu' = [z > 0], relu6' = [0 < z < 6]) and softplus's threshold 20
// (z > 20: z).  gelu / gelu_tf are torch's tanh approximation; silut
// (custom_silu) switches at thr to tanh(slope (z - thr)) + cst with slope
// and cst silu'(thr) and silu(thr) (SiLUT, utils.py:150-171).  Codes are
// DPK_ACT_* of include/dpk.h.
struct ActSpec {
  int id;
  double thr, slope, cst;
};

template <typename T>
__device__ __forceinline__ void act_fd(const ActSpec& a, T z, T& f, T& d) {
  switch (a.id) {
    case DPK_ACT_TANH: {
      const T t = tanh(z);
      f = t;
      d = T(1) - t * t;
      break;
    }
    case DPK_ACT_RELU:
      f = z > T(0) ? z : T(0);
      d = z > T(0) ? T(1) : T(0);
      break;
    case DPK_ACT_RELU6:
      f = z > T(0) ? (z < T(6) ? z : T(6)) : T(0);
      d = (z > T(0) && z < T(6)) ? T(1) : T(0);
      break;
    case DPK_ACT_SOFTPLUS:
      if (z > T(20)) {
        f = z;
        d = T(1);
      } else {
        f = log1p(exp(z));
        d = T(1) / (T(1) + exp(-z));
      }
      break;
    case DPK_ACT_SIGMOID: {
      const T s = T(1) / (T(1) + exp(-z));
      f = s;
      d = s * (T(1) - s);
      break;
    }
    case DPK_ACT_SILU: {
      const T s = T(1) / (T(1) + exp(-z));
      f = z * s;
      d = s + z * s * (T(1) - s);
      break;
    }
    case DPK_ACT_GELU: {
      const T k0 = T(0.7978845608028654), k1 = T(0.044715);
      const T t = tanh(k0 * (z + k1 * z * z * z));
      f = T(0.5) * z * (T(1) + t);
      d = T(0.5) * (T(1) + t) + T(0.5) * z * (T(1) - t * t) * k0 * (T(1) + T(3) * k1 * z * z);
      break;
    }
    case DPK_ACT_SILUT:
      if (z >= T(a.thr)) {
        const T t = tanh(T(a.slope) * (z - T(a.thr)));
        f = t + T(a.cst);
        d = T(a.slope) * (T(1) - t * t);
      } else {
        const T s = T(1) / (T(1) + exp(-z));
        f = z * s;
        d = s + z * s * (T(1) - s);
      }
      break;
    default:  // DPK_ACT_LINEAR
      f = z;
      d = T(1);
      break;
  }
}

// Host side: the spec of an activation code (silut's slope and constant
// from its threshold); false for an unknown code.
inline bool act_spec(int id, double thr, ActSpec* out) {
  if (id < DPK_ACT_TANH || id > DPK_ACT_SILUT) return false;
  const double s = 1.0 / (1.0 + std::exp(-thr));
  *out = ActSpec{id, thr, s + thr * s * (1.0 - s), thr * s};
  return true;
}

// Wave-level sum (all 64 lanes receive the result).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// One past the last real neighbour (index >= 0) among slots [b0, b1) of a
// neighbour-list row (b0 when there is none); the whole wavefront takes part.
__device__ __forceinline__ int block_end(const int* __restrict__ row, int b0, int b1, int lane) {
  int end = b0;
  for (int q0 = b0; q0 < b1; q0 += 64) {
    const int q = q0 + lane;
    const unsigned long long m = __ballot(q < b1 && row[q] >= 0);
    if (m != 0) end = q0 + 64 - __clzll(m);
  }
  return end;
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

}  // namespace dpk
