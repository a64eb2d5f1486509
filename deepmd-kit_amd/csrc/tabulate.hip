// Tabulated (compressed) se_a embedding: tabulate_fusion_se_a forward, grad,
// grad_grad.
//
// Semantics (reference): source/lib/src/tabulate.cc:45-73 (locate_xx: two
// strides + linear extrapolation), :122-158 (quintic and its derivative),
// :162-232 (forward with the is_sorted padding fold: the first slot whose x
// equals the last slot's x and whose em[1:4] are zero stands for all the
// remaining slots, weight nnei - j), :234-341 (grad), :343-447 (grad_grad).
//
// Layout: one workgroup per atom.  Forward / grad_grad put the L channels on
// the threads and walk the neighbours (the fold decision is uniform across the
// workgroup); grad puts the neighbours on the threads and walks the channels,
// with dy for the atom staged in LDS.  The table row for a neighbour's
// interval is contiguous (6 coefficients per channel), so channel-parallel
// reads are unit-stride.
#include "dpk_common.h"

namespace dpk {
namespace {

template <typename T>
struct TabInfo {
  T lower, upper, vmax, stride0, stride1;
  int first;    // int((upper - lower) / stride0)
  int tail;     // table index of the high tail
  int last;     // nspline - 1: interval indices are clamped to the table
};

template <typename T>
TabInfo<T> make_info(const T* h, int nspline) {
  TabInfo<T> ti;
  ti.lower = h[0];
  ti.upper = h[1];
  ti.vmax = h[2];
  ti.stride0 = h[3];
  ti.stride1 = h[4];
  ti.first = (int)((ti.upper - ti.lower) / ti.stride0);
  T bxx = std::nextafter(ti.vmax, ti.lower);
  ti.tail = ti.first + (int)((bxx - ti.upper) / ti.stride1);
  ti.last = nspline - 1;
  return ti;
}

template <typename T>
__device__ __forceinline__ void locate(const TabInfo<T>& ti, T xx, int& idx, T& x, T& ext) {
  ext = T(0);
  if (xx < ti.lower) {
    idx = 0;
    x = T(0);
    ext = xx - ti.lower;
  } else if (xx < ti.upper) {
    idx = (int)((xx - ti.lower) / ti.stride0);
    x = xx - ((T)idx * ti.stride0 + ti.lower);
  } else if (xx < ti.vmax) {
    idx = ti.first + (int)((xx - ti.upper) / ti.stride1);
    x = xx - ((T)(idx - ti.first) * ti.stride1 + ti.upper);
  } else {
    idx = ti.tail;
    x = ti.vmax - ((T)(idx - ti.first) * ti.stride1 + ti.upper);
    ext = xx - ti.vmax;
  }
  idx = idx < 0 ? 0 : (idx > ti.last ? ti.last : idx);
}

// value (with extrapolation) and first derivative of channel k's piece
template <typename T>
__device__ __forceinline__ void piece(const T* __restrict__ c, T x, T ext, T& val, T& grd) {
  const T a0 = c[0], a1 = c[1], a2 = c[2], a3 = c[3], a4 = c[4], a5 = c[5];
  grd = a1 + (T(2) * a2 + (T(3) * a3 + (T(4) * a4 + T(5) * a5 * x) * x) * x) * x;
  val = a0 + (a1 + (a2 + (a3 + (a4 + a5 * x) * x) * x) * x) * x + grd * ext;
}

template <typename T>
__device__ __forceinline__ bool fold_hit(const T* __restrict__ em_x, const T* __restrict__ em,
                                         int nnei, int nd, int jj, T last) {
  return em_x[jj] == last && em[jj * nd + 1] == T(0) && em[jj * nd + 2] == T(0) &&
         em[jj * nd + 3] == T(0);
}

template <typename T, int ND>
__global__ void k_tab_fwd(T* __restrict__ out, const T* __restrict__ table, TabInfo<T> ti,
                          const T* __restrict__ em_x, const T* __restrict__ em, int nnei, int L) {
  const int ii = blockIdx.x;
  const T* ex = em_x + (size_t)ii * nnei;
  const T* e = em + (size_t)ii * nnei * ND;
  const T last = ex[nnei - 1];
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    T acc[ND];
    for (int m = 0; m < ND; ++m) acc[m] = T(0);
    for (int jj = 0; jj < nnei; ++jj) {
      const bool unloop = fold_hit<T>(ex, e, nnei, ND, jj, last);
      int idx;
      T x, ext, val, grd;
      locate<T>(ti, ex[jj], idx, x, ext);
      piece<T>(table + (size_t)idx * L * 6 + 6 * k, x, ext, val, grd);
      const T sc = unloop ? T(nnei - jj) * val : val;
      for (int m = 0; m < ND; ++m) acc[m] += sc * e[jj * ND + m];
      if (unloop) break;
    }
    for (int m = 0; m < ND; ++m) out[((size_t)ii * ND + m) * L + k] = acc[m];
  }
}

template <typename T, int ND>
__global__ void k_tab_grad(T* __restrict__ dx_out, T* __restrict__ dem_out,
                           const T* __restrict__ table, TabInfo<T> ti,
                           const T* __restrict__ em_x, const T* __restrict__ em,
                           const T* __restrict__ dy, int nnei, int L) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* dys = reinterpret_cast<T*>(smem);  // [ND][L]
  __shared__ int first;
  const int ii = blockIdx.x;
  const T* ex = em_x + (size_t)ii * nnei;
  const T* e = em + (size_t)ii * nnei * ND;
  const T last = ex[nnei - 1];
  for (int q = threadIdx.x; q < ND * L; q += blockDim.x) dys[q] = dy[(size_t)ii * ND * L + q];
  if (threadIdx.x == 0) first = nnei;
  __syncthreads();
  for (int jj = threadIdx.x; jj < nnei; jj += blockDim.x)
    if (fold_hit<T>(ex, e, nnei, ND, jj, last)) atomicMin(&first, jj);
  __syncthreads();
  const int f = first;
  for (int jj = threadIdx.x; jj < nnei; jj += blockDim.x) {
    T gx = T(0), gm[ND];
    for (int m = 0; m < ND; ++m) gm[m] = T(0);
    if (jj <= f) {
      T ll[ND];
      for (int m = 0; m < ND; ++m) ll[m] = e[jj * ND + m];
      int idx;
      T x, ext;
      locate<T>(ti, ex[jj], idx, x, ext);
      const T* row = table + (size_t)idx * L * 6;
      for (int k = 0; k < L; ++k) {
        T val, grd;
        piece<T>(row + 6 * k, x, ext, val, grd);
        T dot = T(0);
        for (int m = 0; m < ND; ++m) dot += ll[m] * dys[m * L + k];
        gx += grd * dot;
        for (int m = 0; m < ND; ++m) gm[m] += val * dys[m * L + k];
      }
      const T w = jj == f ? T(nnei - jj) : T(1);
      gx *= w;
      for (int m = 0; m < ND; ++m) gm[m] *= w;
    }
    dx_out[(size_t)ii * nnei + jj] = gx;
    for (int m = 0; m < ND; ++m) dem_out[((size_t)ii * nnei + jj) * ND + m] = gm[m];
  }
}

template <typename T, int ND>
__global__ void k_tab_gg(T* __restrict__ out, const T* __restrict__ table, TabInfo<T> ti,
                         const T* __restrict__ em_x, const T* __restrict__ em,
                         const T* __restrict__ dzx, const T* __restrict__ dzem, int nnei, int L) {
  const int ii = blockIdx.x;
  const T* ex = em_x + (size_t)ii * nnei;
  const T* e = em + (size_t)ii * nnei * ND;
  const T* hx = dzx + (size_t)ii * nnei;
  const T* hm = dzem + (size_t)ii * nnei * ND;
  const T last = ex[nnei - 1];
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    T acc[ND];
    for (int m = 0; m < ND; ++m) acc[m] = T(0);
    for (int jj = 0; jj < nnei; ++jj) {
      const bool unloop = fold_hit<T>(ex, e, nnei, ND, jj, last);
      int idx;
      T x, ext, val, grd;
      locate<T>(ti, ex[jj], idx, x, ext);
      piece<T>(table + (size_t)idx * L * 6 + 6 * k, x, ext, val, grd);
      const T sc = unloop ? T(nnei - jj) : T(1);
      const T gx = hx[jj] * grd;
      for (int m = 0; m < ND; ++m) acc[m] += sc * (val * hm[jj * ND + m] + gx * e[jj * ND + m]);
      if (unloop) break;
    }
    for (int m = 0; m < ND; ++m) out[((size_t)ii * ND + m) * L + k] = acc[m];
  }
}

__global__ void k_zero_any(unsigned char* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int zero_bytes(void* p, size_t n, hipStream_t st) {
  if (n == 0) return DPK_OK;
  hipLaunchKernelGGL(k_zero_any, dim3(cdiv((long long)n, 256)), dim3(256), 0, st,
                     static_cast<unsigned char*>(p), n);
  DPK_LAUNCHED();
  return DPK_OK;
}

int block_for(int L) { return L >= 256 ? 256 : (L > 128 ? 256 : (L > 64 ? 128 : 64)); }

template <typename T>
int check(const T* info, int nloc, int nnei, int L, int nd, int nspline) {
  DPK_REQUIRE(nspline > 0, DPK_ERR_ARG, "tabulate: empty table");
  DPK_REQUIRE(info != nullptr, DPK_ERR_ARG, "tabulate: table_info must be a host array");
  DPK_REQUIRE(nloc >= 0 && nnei >= 0 && L > 0, DPK_ERR_ARG, "tabulate: bad sizes");
  DPK_REQUIRE(nd == 4 || nd == 9 || nd == 16 || nd == 25, DPK_ERR_ARG,
              "The environment basis dimension must be 4, 9, 16, or 25, got %d", nd);
  DPK_REQUIRE(info[3] > T(0) && info[4] > T(0), DPK_ERR_ARG, "tabulate: strides must be > 0");
  return DPK_OK;
}

template <typename T>
int tab_fwd(T* out, const T* table, int nspline, const T* info, const T* em_x, const T* em, int nloc,
            int nnei, int L, int nd, dpk_stream_t stream) {
  clear_error();
  int rc = check<T>(info, nloc, nnei, L, nd, nspline);
  if (rc) return rc;
  hipStream_t st = as_stream(stream);
  if (nloc == 0) return DPK_OK;
  if (nnei == 0) return zero_bytes(out, sizeof(T) * (size_t)nloc * nd * L, st);
  TabInfo<T> ti = make_info<T>(info, nspline);
  dim3 g(nloc), b(block_for(L));
  switch (nd) {
    case 4: hipLaunchKernelGGL((k_tab_fwd<T, 4>), g, b, 0, st, out, table, ti, em_x, em, nnei, L); break;
    case 9: hipLaunchKernelGGL((k_tab_fwd<T, 9>), g, b, 0, st, out, table, ti, em_x, em, nnei, L); break;
    case 16: hipLaunchKernelGGL((k_tab_fwd<T, 16>), g, b, 0, st, out, table, ti, em_x, em, nnei, L); break;
    default: hipLaunchKernelGGL((k_tab_fwd<T, 25>), g, b, 0, st, out, table, ti, em_x, em, nnei, L); break;
  }
  DPK_LAUNCHED();
  return DPK_OK;
}

template <typename T>
int tab_grad(T* dx, T* dem, const T* table, int nspline, const T* info, const T* em_x, const T* em,
             const T* dy, int nloc, int nnei, int L, int nd, dpk_stream_t stream) {
  clear_error();
  int rc = check<T>(info, nloc, nnei, L, nd, nspline);
  if (rc) return rc;
  if (nloc == 0 || nnei == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  TabInfo<T> ti = make_info<T>(info, nspline);
  size_t lds = sizeof(T) * (size_t)nd * L;
  DPK_REQUIRE(lds <= 60 * 1024, DPK_ERR_UNSUPPORTED, "tabulate grad: nd*L too large for LDS");
  dim3 g(nloc), b(nnei >= 128 ? 128 : 64);
  switch (nd) {
    case 4: hipLaunchKernelGGL((k_tab_grad<T, 4>), g, b, lds, st, dx, dem, table, ti, em_x, em, dy, nnei, L); break;
    case 9: hipLaunchKernelGGL((k_tab_grad<T, 9>), g, b, lds, st, dx, dem, table, ti, em_x, em, dy, nnei, L); break;
    case 16: hipLaunchKernelGGL((k_tab_grad<T, 16>), g, b, lds, st, dx, dem, table, ti, em_x, em, dy, nnei, L); break;
    default: hipLaunchKernelGGL((k_tab_grad<T, 25>), g, b, lds, st, dx, dem, table, ti, em_x, em, dy, nnei, L); break;
  }
  DPK_LAUNCHED();
  return DPK_OK;
}

template <typename T>
int tab_gg(T* out, const T* table, int nspline, const T* info, const T* em_x, const T* em, const T* dzx,
           const T* dzem, int nloc, int nnei, int L, int nd, dpk_stream_t stream) {
  clear_error();
  int rc = check<T>(info, nloc, nnei, L, nd, nspline);
  if (rc) return rc;
  hipStream_t st = as_stream(stream);
  if (nloc == 0) return DPK_OK;
  if (nnei == 0) return zero_bytes(out, sizeof(T) * (size_t)nloc * nd * L, st);
  TabInfo<T> ti = make_info<T>(info, nspline);
  dim3 g(nloc), b(block_for(L));
  switch (nd) {
    case 4: hipLaunchKernelGGL((k_tab_gg<T, 4>), g, b, 0, st, out, table, ti, em_x, em, dzx, dzem, nnei, L); break;
    case 9: hipLaunchKernelGGL((k_tab_gg<T, 9>), g, b, 0, st, out, table, ti, em_x, em, dzx, dzem, nnei, L); break;
    case 16: hipLaunchKernelGGL((k_tab_gg<T, 16>), g, b, 0, st, out, table, ti, em_x, em, dzx, dzem, nnei, L); break;
    default: hipLaunchKernelGGL((k_tab_gg<T, 25>), g, b, 0, st, out, table, ti, em_x, em, dzx, dzem, nnei, L); break;
  }
  DPK_LAUNCHED();
  return DPK_OK;
}

// ---------------------------------------------------------------------------
// Compressed embedding inside the fused model: the per-type-block loop of
// deepmd/pt/model/descriptor/se_a.py:795-843 (one tabulate_fusion_se_a call
// per (centre type, neighbour type) table, summed, divided by nnei) over the
// env matrix em [natoms, nnei, 4] in place: em_x of a slot is em[slot][0] and
// the padding fold applies per type block (the block's last slot).  Tables of
// all nets are concatenated; their infos travel by value.
constexpr int kMaxTabNets = 16;
constexpr int kMaxTabTypes = 16;

template <typename T>
struct TabSet {
  TabInfo<T> ti[kMaxTabNets];
  long long off[kMaxTabNets];  // element offset of net k's table
  int ntypes;
  int sec[kMaxTabTypes + 1];
};

template <typename T>
__device__ __forceinline__ bool slot_fold(const T* r, T last) {
  return r[0] == last && r[1] == T(0) && r[2] == T(0) && r[3] == T(0);
}

template <typename T>
__global__ void __launch_bounds__(128) k_tabemb_fwd(T* __restrict__ xyz, const T* __restrict__ em,
                                                    const int* __restrict__ atype,
                                                    const T* __restrict__ tables, TabSet<T> ts,
                                                    const int* __restrict__ net_of, int L, int nnei,
                                                    T inv_nnei) {
  const int atom = blockIdx.x;
  int tc = atype[atom];
  tc = tc < 0 ? 0 : tc;
  const T* e = em + (size_t)atom * nnei * 4;
  for (int k = threadIdx.x; k < L; k += blockDim.x) {
    T acc[4] = {T(0), T(0), T(0), T(0)};
    for (int t = 0; t < ts.ntypes; ++t) {
      const int b0 = ts.sec[t], nb = ts.sec[t + 1] - b0;
      if (nb <= 0) continue;
      const int net = net_of[tc * ts.ntypes + t];
      const TabInfo<T>& ti = ts.ti[net];
      const T* tab = tables + ts.off[net];
      const T last = e[(size_t)(b0 + nb - 1) * 4];
      for (int jj = 0; jj < nb; ++jj) {
        const T* r = e + (size_t)(b0 + jj) * 4;
        const bool unloop = slot_fold<T>(r, last);
        int idx;
        T x, ext, val, grd;
        locate<T>(ti, r[0], idx, x, ext);
        piece<T>(tab + (size_t)idx * L * 6 + 6 * k, x, ext, val, grd);
        const T sc = unloop ? T(nb - jj) * val : val;
        for (int m = 0; m < 4; ++m) acc[m] += sc * r[m];
        if (unloop) break;
      }
    }
    for (int m = 0; m < 4; ++m) xyz[((size_t)atom * 4 + m) * L + k] = acc[m] * inv_nnei;
  }
}

template <typename T>
__global__ void __launch_bounds__(128) k_tabemb_bwd(T* __restrict__ nd, const T* __restrict__ em,
                                                    const T* __restrict__ gxyz,
                                                    const int* __restrict__ atype,
                                                    const T* __restrict__ tables, TabSet<T> ts,
                                                    const int* __restrict__ net_of, int L, int nnei,
                                                    T inv_nnei) {
  extern __shared__ __align__(16) unsigned char smem[];
  T* dys = reinterpret_cast<T*>(smem);  // [4][L] = dE/dxyz / nnei
  __shared__ int first[kMaxTabTypes];
  const int atom = blockIdx.x;
  int tc = atype[atom];
  tc = tc < 0 ? 0 : tc;
  const T* e = em + (size_t)atom * nnei * 4;
  for (int q = threadIdx.x; q < 4 * L; q += blockDim.x)
    dys[q] = gxyz[(size_t)atom * 4 * L + q] * inv_nnei;
  for (int t = threadIdx.x; t < ts.ntypes; t += blockDim.x) first[t] = ts.sec[t + 1] - ts.sec[t];
  __syncthreads();
  for (int t = 0; t < ts.ntypes; ++t) {
    const int b0 = ts.sec[t], nb = ts.sec[t + 1] - b0;
    if (nb <= 0) continue;
    const T last = e[(size_t)(b0 + nb - 1) * 4];
    for (int jj = threadIdx.x; jj < nb; jj += blockDim.x)
      if (slot_fold<T>(e + (size_t)(b0 + jj) * 4, last)) atomicMin(&first[t], jj);
  }
  __syncthreads();
  for (int t = 0; t < ts.ntypes; ++t) {
    const int b0 = ts.sec[t], nb = ts.sec[t + 1] - b0;
    if (nb <= 0) continue;
    const int net = net_of[tc * ts.ntypes + t];
    const TabInfo<T>& ti = ts.ti[net];
    const T* tab = tables + ts.off[net];
    const int f = first[t];
    for (int jj = threadIdx.x; jj < nb; jj += blockDim.x) {
      const T* r = e + (size_t)(b0 + jj) * 4;
      T gx = T(0), gm[4] = {T(0), T(0), T(0), T(0)};
      if (jj <= f) {
        int idx;
        T x, ext;
        locate<T>(ti, r[0], idx, x, ext);
        const T* row = tab + (size_t)idx * L * 6;
        const T r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
        for (int k = 0; k < L; ++k) {
          T val, grd;
          piece<T>(row + 6 * k, x, ext, val, grd);
          const T d0 = dys[k], d1 = dys[L + k], d2 = dys[2 * L + k], d3 = dys[3 * L + k];
          gx += grd * (r0 * d0 + r1 * d1 + r2 * d2 + r3 * d3);
          gm[0] += val * d0;
          gm[1] += val * d1;
          gm[2] += val * d2;
          gm[3] += val * d3;
        }
        const T w = jj == f ? T(nb - jj) : T(1);
        gx *= w;
        for (int m = 0; m < 4; ++m) gm[m] *= w;
      }
      T* o = nd + ((size_t)atom * nnei + b0 + jj) * 4;
      o[0] = gm[0] + gx;  // em_x is em[..., 0]
      o[1] = gm[1];
      o[2] = gm[2];
      o[3] = gm[3];
    }
  }
}

template <typename T>
int tabemb(bool fwd, void* out, const void* em, const void* gxyz, const int* atype,
           const void* tables, const long long* off, const int* nspline, const double* info,
           const int* net_of, int ntypes, int nnets, const int* sec_host, int L, int natoms,
           int nnei, hipStream_t st) {
  DPK_REQUIRE(ntypes > 0 && ntypes <= kMaxTabTypes && nnets > 0 && nnets <= kMaxTabNets,
              DPK_ERR_UNSUPPORTED, "tab_embed: ntypes=%d nnets=%d (<= %d tables)", ntypes, nnets,
              kMaxTabNets);
  DPK_REQUIRE(L > 0 && nnei > 0 && natoms >= 0, DPK_ERR_ARG, "tab_embed: bad sizes");
  TabSet<T> ts;
  ts.ntypes = ntypes;
  for (int t = 0; t <= ntypes; ++t) ts.sec[t] = sec_host[t];
  DPK_REQUIRE(ts.sec[ntypes] == nnei, DPK_ERR_ARG, "tab_embed: sum(sel)=%d != nnei=%d",
              ts.sec[ntypes], nnei);
  for (int k = 0; k < nnets; ++k) {
    T h[6];
    for (int q = 0; q < 6; ++q) h[q] = (T)info[6 * k + q];
    DPK_REQUIRE(nspline[k] > 0 && h[3] > T(0) && h[4] > T(0), DPK_ERR_ARG,
                "tab_embed: table %d is empty or has a non-positive stride", k);
    ts.ti[k] = make_info<T>(h, nspline[k]);
    ts.off[k] = off[k];
  }
  if (natoms == 0) return DPK_OK;
  const T inv = T(1) / T(nnei);
  if (fwd) {
    hipLaunchKernelGGL((k_tabemb_fwd<T>), dim3(natoms), dim3(128), 0, st, static_cast<T*>(out),
                       static_cast<const T*>(em), atype, static_cast<const T*>(tables), ts, net_of,
                       L, nnei, inv);
  } else {
    const size_t lds = sizeof(T) * 4 * (size_t)L;
    DPK_REQUIRE(lds <= 32 * 1024, DPK_ERR_UNSUPPORTED, "tab_embed: L=%d too large", L);
    hipLaunchKernelGGL((k_tabemb_bwd<T>), dim3(natoms), dim3(128), lds, st, static_cast<T*>(out),
                       static_cast<const T*>(em), static_cast<const T*>(gxyz), atype,
                       static_cast<const T*>(tables), ts, net_of, L, nnei, inv);
  }
  DPK_LAUNCHED();
  return DPK_OK;
}

int tabemb_common(bool fwd, void* out, int prec, const void* em, const void* gxyz,
                  const int* atype, const void* tables, const long long* off, const int* nspline,
                  const double* info, const int* net_of, int ntypes, int nnets,
                  const int* sec_host, int L, int natoms, int nnei, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(off && nspline && info && sec_host, DPK_ERR_ARG,
              "tab_embed: table offsets, sizes, infos and sections are host arrays");
  hipStream_t st = as_stream(stream);
  if (prec == DPK_PREC_F64)
    return tabemb<double>(fwd, out, em, gxyz, atype, tables, off, nspline, info, net_of, ntypes,
                          nnets, sec_host, L, natoms, nnei, st);
  if (prec == DPK_PREC_F32)
    return tabemb<float>(fwd, out, em, gxyz, atype, tables, off, nspline, info, net_of, ntypes,
                         nnets, sec_host, L, natoms, nnei, st);
  set_error("tab_embed: precision %d unsupported (FP64 or FP32 tables)", prec);
  return DPK_ERR_UNSUPPORTED;
}

}  // namespace
}  // namespace dpk

using namespace dpk;

extern "C" int dpk_tab_embed_fwd(void* xyz, int prec, const void* em, const int* atype,
                                 const void* tables, const long long* table_off,
                                 const int* nspline, const double* info, const int* net_of,
                                 int ntypes, int nnets, const int* sec_host, int L, int natoms,
                                 int nnei, dpk_stream_t stream) {
  return tabemb_common(true, xyz, prec, em, nullptr, atype, tables, table_off, nspline, info,
                       net_of, ntypes, nnets, sec_host, L, natoms, nnei, stream);
}

extern "C" int dpk_tab_embed_bwd(void* net_deriv, int prec, const void* em, const void* gxyz,
                                 const int* atype, const void* tables, const long long* table_off,
                                 const int* nspline, const double* info, const int* net_of,
                                 int ntypes, int nnets, const int* sec_host, int L, int natoms,
                                 int nnei, dpk_stream_t stream) {
  return tabemb_common(false, net_deriv, prec, em, gxyz, atype, tables, table_off, nspline, info,
                       net_of, ntypes, nnets, sec_host, L, natoms, nnei, stream);
}

extern "C" int dpk_tabulate_fusion_se_a_f64(double* out, const double* table, int nspline,
                                            const double* info, const double* em_x,
                                            const double* em, int nloc, int nnei, int L, int nd,
                                            dpk_stream_t s) {
  return tab_fwd<double>(out, table, nspline, info, em_x, em, nloc, nnei, L, nd, s);
}
extern "C" int dpk_tabulate_fusion_se_a_f32(float* out, const float* table, int nspline,
                                            const float* info,
                                            const float* em_x, const float* em, int nloc,
                                            int nnei, int L, int nd, dpk_stream_t s) {
  return tab_fwd<float>(out, table, nspline, info, em_x, em, nloc, nnei, L, nd, s);
}
extern "C" int dpk_tabulate_fusion_se_a_grad_f64(double* dx, double* dem, const double* table,
                                                 int nspline, const double* info, const double* em_x,
                                                 const double* em, const double* dy, int nloc,
                                                 int nnei, int L, int nd, dpk_stream_t s) {
  return tab_grad<double>(dx, dem, table, nspline, info, em_x, em, dy, nloc, nnei, L, nd, s);
}
extern "C" int dpk_tabulate_fusion_se_a_grad_f32(float* dx, float* dem, const float* table,
                                                 int nspline, const float* info, const float* em_x,
                                                 const float* em, const float* dy, int nloc,
                                                 int nnei, int L, int nd, dpk_stream_t s) {
  return tab_grad<float>(dx, dem, table, nspline, info, em_x, em, dy, nloc, nnei, L, nd, s);
}
extern "C" int dpk_tabulate_fusion_se_a_grad_grad_f64(double* out, const double* table,
                                                      int nspline, const double* info, const double* em_x,
                                                      const double* em, const double* dzx,
                                                      const double* dzem, int nloc, int nnei,
                                                      int L, int nd, dpk_stream_t s) {
  return tab_gg<double>(out, table, nspline, info, em_x, em, dzx, dzem, nloc, nnei, L, nd, s);
}
extern "C" int dpk_tabulate_fusion_se_a_grad_grad_f32(float* out, const float* table,
                                                      int nspline, const float* info, const float* em_x,
                                                      const float* em, const float* dzx,
                                                      const float* dzem, int nloc, int nnei,
                                                      int L, int nd, dpk_stream_t s) {
  return tab_gg<float>(out, table, nspline, info, em_x, em, dzx, dzem, nloc, nnei, L, nd, s);
}
