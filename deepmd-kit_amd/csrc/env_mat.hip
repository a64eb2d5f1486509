// Smooth environment matrix of se_e2_a.
//
// Per slot (centre i, neighbour j): d = r_j - r_i, r = |d|,
//   sw = switch(r; rcut_smth, rcut), R = sw * (1/r, x/r^2, y/r^2, z/r^2),
//   em = (R - avg[t_i]) / std[t_i];  padded slots (nlist < 0) give R = 0.
// Reference: deepmd/pt/model/descriptor/env_mat.py:18-119,
// deepmd/pt/utils/preprocess.py:9-18 (value), source/lib/src/env_mat.cc:123-221
// (derivative w.r.t. the centre), source/lib/src/prod_env_mat.cc:95-119
// (normalisation, em_deriv / std).
//
// Two forms: the op form (explicit extended coordinates, em + em_deriv + rij,
// the torch op `prod_env_mat_a`) and the fused form (implicit periodic images,
// em only, feeding the embedding kernels).  One thread per slot; every slot
// writes its outputs with unit-stride rows (coalesced across the wavefront).
#include "dpk_common.h"

namespace dpk {
namespace {

// Value and d/d r_centre of the four components (C++ env_mat.cc:162-211 form;
// with u = r + protection, R = sw (1/u, d/u^2), env_mat.py:38-45).
template <typename T>
__device__ __forceinline__ void env_value_deriv(T x, T y, T z, T r, T rmin, T rmax, T prot,
                                                T em[4], T dem[12]) {
  const T inr = T(1) / r;
  const T iu = prot == T(0) ? inr : T(1) / (r + prot);
  T sw, dsw;
  spline5<T>(r, rmin, rmax, sw, dsw);
  const T d[3] = {x, y, z};
  // dem[3 c + a] = d R_c / d r_centre[a]: the contraction with each unit g
  for (int c = 0; c < 4; ++c) {
    const T gu[4] = {T(c == 0), T(c == 1), T(c == 2), T(c == 3)};
    T f[3];
    env_centre_deriv<T>(d, inr, iu, sw, dsw, gu, f);
    for (int a = 0; a < 3; ++a) dem[3 * c + a] = f[a];
  }
  const T iu2 = iu * iu;
  em[0] = iu * sw;
  em[1] = x * iu2 * sw;
  em[2] = y * iu2 * sw;
  em[3] = z * iu2 * sw;
}

template <typename T>
__global__ void k_env_op(T* __restrict__ em, T* __restrict__ em_deriv, T* __restrict__ rij,
                         const T* __restrict__ coord, const int* __restrict__ atype,
                         const int* __restrict__ nlist, const T* __restrict__ avg,
                         const T* __restrict__ stdv, int nframes, int nloc, int nall, int nnei,
                         int ntypes, T rcut, T rcut_smth, T prot) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long total = (long long)nframes * nloc * nnei;
  if (idx >= total) return;
  const int jj = (int)(idx % nnei);
  const long long row = idx / nnei;  // frame * nloc + i
  const int ff = (int)(row / nloc);
  const int ii = (int)(row % nloc);
  const int* at = atype + (size_t)ff * nall;
  const T* cf = coord + (size_t)ff * nall * 3;
  // a centre type outside [0, ntypes) has no statistics row: zero output,
  // like a virtual atom; neighbour indices outside [0, nall) are empty slots
  const int ti = at[ii] < ntypes ? at[ii] : -1;
  T e4[4] = {0, 0, 0, 0}, d12[12], r3[3] = {0, 0, 0};
  for (int k = 0; k < 12; ++k) d12[k] = 0;
  const int j = nlist[idx];
  if (ti >= 0 && j >= 0 && j < nall) {
    T x = cf[3 * j] - cf[3 * ii], y = cf[3 * j + 1] - cf[3 * ii + 1],
      z = cf[3 * j + 2] - cf[3 * ii + 2];
    r3[0] = x;
    r3[1] = y;
    r3[2] = z;
    T r = sqrt(x * x + y * y + z * z);
    env_value_deriv<T>(x, y, z, r, rcut_smth, rcut, prot, e4, d12);
  }
  T* eo = em + idx * 4;
  T* dout = em_deriv + idx * 12;
  if (ti >= 0) {
    const T* a = avg + (size_t)ti * nnei * 4 + jj * 4;
    const T* s = stdv + (size_t)ti * nnei * 4 + jj * 4;
    for (int c = 0; c < 4; ++c) {
      eo[c] = (e4[c] - a[c]) / s[c];
      for (int k = 0; k < 3; ++k) dout[3 * c + k] = d12[3 * c + k] / s[c];
    }
  } else {
    for (int c = 0; c < 4; ++c) eo[c] = T(0);
    for (int k = 0; k < 12; ++k) dout[k] = T(0);
  }
  T* ro = rij + idx * 3;
  ro[0] = r3[0];
  ro[1] = r3[1];
  ro[2] = r3[2];
}

// Fused form: FP64 geometry with implicit images, em stored in ET (float/double).
// One wavefront per centre atom, lanes over its slots (the atom's frame,
// type, position and image decoder are wave-uniform: no per-slot division of
// the flat index).  With ET = float the divisions by r and by std are
// reciprocal-multiplies (within 1 ulp of FP64 division, far below the FP32
// rounding of em); ET = double keeps the exact divisions of the reference.
template <typename ET>
__global__ void __launch_bounds__(256) k_env_fused(
    ET* __restrict__ em, const double* __restrict__ wrapped, const int* __restrict__ atype,
    const int* __restrict__ frame_of, const dpk_frame* __restrict__ frames,
    const int* __restrict__ nlist, const double* __restrict__ shift_vec,
    const double* __restrict__ avg, const double* __restrict__ stdv, int natoms, int nnei,
    double rcut, double rcut_smth, double prot, const int* __restrict__ pair_mask, int ntypes) {
  constexpr bool kExact = sizeof(ET) == 8;
  const int g = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (g >= natoms) return;
  const int traw = atype[g];
  const int ti = traw < 0 ? 0 : traw;
  const int* pm = pair_mask ? pair_mask + (traw < 0 ? ntypes : traw) * (ntypes + 1) : nullptr;
  const dpk_frame& fd = frames[frame_of[g]];
  FastDiv byloc;
  byloc.init((unsigned)fd.nloc);
  const double cx = wrapped[3 * g], cy = wrapped[3 * g + 1], cz = wrapped[3 * g + 2];
  const double* a = avg + (size_t)ti * nnei * 4;
  const double* sd = stdv + (size_t)ti * nnei * 4;
  const int* nl = nlist + (size_t)g * nnei;
  ET* orow = em + (size_t)g * nnei * 4;
  // kBatch slots per lane at a time: their list entries, then all their
  // coordinate gathers, are issued before any arithmetic, so a wavefront has
  // kBatch dependent load chains in flight instead of one
  constexpr int kBatch = 4;
  for (int j0 = 0; j0 < nnei; j0 += kBatch * kWave) {
    int ev[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int jj = j0 + u * kWave + lane;
      ev[u] = jj < nnei ? nl[jj] : -1;
    }
    // a padding slot (e < 0) has R = 0: its row is (0 - avg) / std; an
    // excluded pair's row is zero
    double px[kBatch], py[kBatch], pz[kBatch];
    bool zero[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      px[u] = py[u] = pz[u] = 0.0;
      zero[u] = false;
      const int e = ev[u];
      if (e >= 0) {
        const int s = (int)byloc.div((unsigned)e);
        const int j = fd.atom_off + (e - s * fd.nloc);
        // excluded type pair: the row is zero (se_a.py:811-819, rr * mask); a
        // virtual centre counts as type ntypes, which is never excluded
        zero[u] = pm != nullptr && !pm[atype[j]];
        const double* sv = shift_vec + 3 * (size_t)(fd.img_off + s);
        px[u] = wrapped[3 * j] + sv[0];
        py[u] = wrapped[3 * j + 1] + sv[1];
        pz[u] = wrapped[3 * j + 2] + sv[2];
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int jj = j0 + u * kWave + lane;
      if (jj >= nnei) continue;
      double R[4] = {0.0, 0.0, 0.0, 0.0};
      if (ev[u] >= 0 && !zero[u]) {
        const double x = px[u] - cx;
        const double y = py[u] - cy;
        const double z = pz[u] - cz;
        const double r = norm3_ref(x, y, z);
        double sw, dsw;
        spline5<double>(r, rcut_smth, rcut, sw, dsw);
        const double ru = prot == 0.0 ? r : r + prot;
        const double inv = kExact ? 1.0 / ru : rcp_nr(ru);
        const double inv2 = inv * inv;
        R[0] = inv * sw;
        R[1] = x * inv2 * sw;
        R[2] = y * inv2 * sw;
        R[3] = z * inv2 * sw;
      }
      ET v[4];
      for (int c = 0; c < 4; ++c) {
        const double q = kExact ? (R[c] - a[jj * 4 + c]) / sd[jj * 4 + c]
                                : (R[c] - a[jj * 4 + c]) * rcp_nr(sd[jj * 4 + c]);
        v[c] = zero[u] ? (ET)0 : (ET)q;
      }
      // one 16-byte (float) or two 16-byte (double) stores per slot
      if constexpr (kExact) {
        double2* o = reinterpret_cast<double2*>(orow + jj * 4);
        o[0] = make_double2(v[0], v[1]);
        o[1] = make_double2(v[2], v[3]);
      } else {
        *reinterpret_cast<float4*>(orow + jj * 4) = make_float4(v[0], v[1], v[2], v[3]);
      }
    }
  }
}

}  // namespace
}  // namespace dpk

using namespace dpk;

template <typename T>
static int env_op(T* em, T* em_deriv, T* rij, const T* coord, const int* atype, const int* nlist,
                  const T* avg, const T* stdv, int nframes, int nloc, int nall, int nnei,
                  int ntypes, double rcut, double rcut_smth, double protection,
                  dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(protection >= 0.0, DPK_ERR_ARG, "prod_env_mat_a: protection < 0");
  DPK_REQUIRE(nframes >= 0 && nloc >= 0 && nall >= nloc && nnei > 0 && ntypes > 0, DPK_ERR_ARG,
              "prod_env_mat_a: bad sizes nf=%d nloc=%d nall=%d nnei=%d ntypes=%d", nframes,
              nloc, nall, nnei, ntypes);
  DPK_REQUIRE(rcut > rcut_smth, DPK_ERR_ARG, "prod_env_mat_a: rcut_smth must be < rcut");
  long long total = (long long)nframes * nloc * nnei;
  if (total == 0) return DPK_OK;
  hipLaunchKernelGGL(k_env_op<T>, dim3(cdiv(total, 256)), dim3(256), 0, as_stream(stream), em,
                     em_deriv, rij, coord, atype, nlist, avg, stdv, nframes, nloc, nall, nnei,
                     ntypes, (T)rcut, (T)rcut_smth, (T)protection);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_prod_env_mat_a_f64(double* em, double* em_deriv, double* rij,
                                      const double* coord, const int* atype, const int* nlist,
                                      const double* avg, const double* std, int nframes,
                                      int nloc, int nall, int nnei, int ntypes, double rcut,
                                      double rcut_smth, double protection,
                                      dpk_stream_t stream) {
  return env_op<double>(em, em_deriv, rij, coord, atype, nlist, avg, std, nframes, nloc, nall,
                        nnei, ntypes, rcut, rcut_smth, protection, stream);
}

extern "C" int dpk_prod_env_mat_a_f32(float* em, float* em_deriv, float* rij,
                                      const float* coord, const int* atype, const int* nlist,
                                      const float* avg, const float* std, int nframes, int nloc,
                                      int nall, int nnei, int ntypes, double rcut,
                                      double rcut_smth, double protection,
                                      dpk_stream_t stream) {
  return env_op<float>(em, em_deriv, rij, coord, atype, nlist, avg, std, nframes, nloc, nall,
                       nnei, ntypes, rcut, rcut_smth, protection, stream);
}

extern "C" int dpk_env_mat_fused(void* em, int prec, const double* coord_wrapped,
                                 const int* atype, const int* frame_of, const dpk_frame* frames,
                                 const int* nlist, const double* shift_vec, const double* avg,
                                 const double* std, int natoms, int nnei, double rcut,
                                 double rcut_smth, double protection, const int* pair_mask,
                                 int ntypes, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && nnei > 0, DPK_ERR_ARG, "env_mat_fused: bad sizes");
  DPK_REQUIRE(protection >= 0.0, DPK_ERR_ARG, "env_mat_fused: protection < 0");
  DPK_REQUIRE(rcut > rcut_smth, DPK_ERR_ARG, "env_mat_fused: rcut_smth must be < rcut");
  long long total = (long long)natoms * nnei;
  if (total == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  if (prec == DPK_PREC_F64) {
    hipLaunchKernelGGL(k_env_fused<double>, dim3(cdiv(natoms, 4)), dim3(256), 0, st,
                       static_cast<double*>(em), coord_wrapped, atype, frame_of, frames, nlist,
                       shift_vec, avg, std, natoms, nnei, rcut, rcut_smth, protection, pair_mask,
                       ntypes);
  } else if (prec == DPK_PREC_F32 || prec == DPK_PREC_BF16 || prec == DPK_PREC_F16 ||
             prec == DPK_PREC_F16X2) {
    hipLaunchKernelGGL(k_env_fused<float>, dim3(cdiv(natoms, 4)), dim3(256), 0, st,
                       static_cast<float*>(em), coord_wrapped, atype, frame_of, frames, nlist,
                       shift_vec, avg, std, natoms, nnei, rcut, rcut_smth, protection, pair_mask,
                       ntypes);
  } else {
    set_error("env_mat_fused: unknown precision %d", prec);
    return DPK_ERR_ARG;
  }
  DPK_LAUNCHED();
  return DPK_OK;
}
