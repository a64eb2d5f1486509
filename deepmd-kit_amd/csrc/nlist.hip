// Cell-list neighbour search + per-type formatting (PT nlist semantics).
//
// Behaviour reproduced (reference files):
//   deepmd/pt/utils/region.py:86-105     normalize_coord (remainder(frac, 1))
//   deepmd/pt/utils/nlist.py:424-513     images: extended index e = s*nloc + j
//   deepmd/pt/utils/nlist.py:49-150      Σsel nearest by FP64 |d|, rr <= rcut
//   deepmd/pt/utils/nlist.py:304-337     stable per-type partition, sel[t] cut
// The reference builds a dense [nloc, nall] distance matrix and runs topk; here
// the atoms of each frame are binned into a fractional cell grid (cell width >=
// rcut, so the 27 (or 2*rng+1 per axis) surrounding cells hold every
// neighbour), images are implicit (shift vector added on the fly), and one
// wavefront per centre atom collects the within-rcut candidates in LDS and
// ranks them by (distance, extended index).
#include "dpk_common.h"

namespace dpk {
namespace {

constexpr int kMaxTypes = 32;

struct SelInfo {
  int ntypes;
  int nnei;
  int sel[kMaxTypes];
  int sec[kMaxTypes + 1];
};

struct Cand {
  double r;
  int e;
  int t;
};

__global__ void k_zero(int* __restrict__ cell_count, int ncells, int* __restrict__ status) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ncells) cell_count[i] = 0;
  if (i == 0) status[0] = 0;
}

// torch.remainder(a, 1.0) for doubles: fmod, then shift negatives into [0, 1).
__device__ __forceinline__ double rem1(double a) {
  double m = fmod(a, 1.0);
  if (m != 0.0 && m < 0.0) m += 1.0;
  return m;
}

__global__ void k_bin(const double* __restrict__ coord, const int* __restrict__ frame_of,
                      const dpk_frame* __restrict__ frames, int natoms,
                      double* __restrict__ wrapped, int* __restrict__ atom_cell,
                      int* __restrict__ atom_slot, int* __restrict__ cell_count) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= natoms) return;
  const dpk_frame& fd = frames[frame_of[g]];
  double x[3] = {coord[3 * g], coord[3 * g + 1], coord[3 * g + 2]};
  double w[3], f[3];
  if (fd.pbc) {
    // frac = x @ rec (row vector times matrix), wrapped = rem1(frac) @ box
    for (int k = 0; k < 3; ++k)
      f[k] = __dadd_rn(__dadd_rn(__dmul_rn(x[0], fd.rec[k]), __dmul_rn(x[1], fd.rec[3 + k])),
                       __dmul_rn(x[2], fd.rec[6 + k]));
    for (int k = 0; k < 3; ++k) f[k] = rem1(f[k]);
    for (int k = 0; k < 3; ++k)
      w[k] = __dadd_rn(__dadd_rn(__dmul_rn(f[0], fd.box[k]), __dmul_rn(f[1], fd.box[3 + k])),
                       __dmul_rn(f[2], fd.box[6 + k]));
  } else {
    for (int k = 0; k < 3; ++k) {
      w[k] = x[k];
      f[k] = (x[k] - fd.origin[k]) * fd.inv_extent[k];
    }
  }
  int c[3];
  for (int k = 0; k < 3; ++k) {
    int ck = (int)floor(f[k] * fd.nc[k]);
    c[k] = ck < 0 ? 0 : (ck >= fd.nc[k] ? fd.nc[k] - 1 : ck);
  }
  int cid = fd.cell_off + (c[0] * fd.nc[1] + c[1]) * fd.nc[2] + c[2];
  wrapped[3 * g] = w[0];
  wrapped[3 * g + 1] = w[1];
  wrapped[3 * g + 2] = w[2];
  atom_cell[g] = cid;
  atom_slot[g] = atomicAdd(cell_count + cid, 1);
}

// Exclusive scan of cell counts by a single 1024-thread block.
__global__ void __launch_bounds__(1024) k_scan(const int* __restrict__ cnt, int n,
                                               int* __restrict__ start) {
  __shared__ int part[1024];
  int t = threadIdx.x;
  int per = (n + 1023) / 1024;
  int b = t * per, e = min(n, b + per);
  int s = 0;
  for (int i = b; i < e; ++i) s += cnt[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = part[t] - s;
  for (int i = b; i < e; ++i) {
    start[i] = run;
    run += cnt[i];
  }
  if (t == 1023) start[n] = part[1023];
}

__global__ void k_scatter(const int* __restrict__ atom_cell, const int* __restrict__ atom_slot,
                          const int* __restrict__ cell_start, int natoms,
                          int* __restrict__ cell_atoms) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= natoms) return;
  cell_atoms[cell_start[atom_cell[g]] + atom_slot[g]] = g;
}

__device__ __forceinline__ int floordiv(int a, int b) {
  int q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

// One wavefront (= one block of 64) per centre atom.
template <int MAXC>
__global__ void __launch_bounds__(64) k_search(
    const double* __restrict__ wrapped, const int* __restrict__ atype,
    const int* __restrict__ frame_of, const dpk_frame* __restrict__ frames,
    const int* __restrict__ atom_cell, const int* __restrict__ cell_start,
    const int* __restrict__ cell_atoms, const int* __restrict__ img_pos,
    const double* __restrict__ shift_vec, double rcut, SelInfo si, int natoms,
    int* __restrict__ nlist, int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  Cand* cand = reinterpret_cast<Cand*>(smem);
  int* row = reinterpret_cast<int*>(smem + sizeof(Cand) * MAXC);
  const int g = blockIdx.x;
  const int lane = threadIdx.x;
  if (g >= natoms) return;
  const int S = si.nnei;
  int* out = nlist + (size_t)g * S;
  const int ti = atype[g];
  if (ti < 0) {  // virtual centre: empty row (nlist.py:193-195)
    for (int s = lane; s < S; s += kWave) out[s] = -1;
    return;
  }
  const dpk_frame& fd = frames[frame_of[g]];
  const double cx = wrapped[3 * g], cy = wrapped[3 * g + 1], cz = wrapped[3 * g + 2];
  const int loc = atom_cell[g] - fd.cell_off;
  const int c2 = loc % fd.nc[2];
  const int c1 = (loc / fd.nc[2]) % fd.nc[1];
  const int c0 = loc / (fd.nc[2] * fd.nc[1]);
  const int w1 = 2 * fd.nb[1] + 1, w2 = 2 * fd.nb[2] + 1;
  int n = 0;
  for (int d0 = -fd.rng[0]; d0 <= fd.rng[0]; ++d0) {
    int u0 = c0 + d0, s0 = 0;
    if (fd.pbc) { s0 = floordiv(u0, fd.nc[0]); u0 -= s0 * fd.nc[0]; }
    else if (u0 < 0 || u0 >= fd.nc[0]) continue;
    for (int d1 = -fd.rng[1]; d1 <= fd.rng[1]; ++d1) {
      int u1 = c1 + d1, s1 = 0;
      if (fd.pbc) { s1 = floordiv(u1, fd.nc[1]); u1 -= s1 * fd.nc[1]; }
      else if (u1 < 0 || u1 >= fd.nc[1]) continue;
      for (int d2 = -fd.rng[2]; d2 <= fd.rng[2]; ++d2) {
        int u2 = c2 + d2, s2 = 0;
        if (fd.pbc) { s2 = floordiv(u2, fd.nc[2]); u2 -= s2 * fd.nc[2]; }
        else if (u2 < 0 || u2 >= fd.nc[2]) continue;
        if (s0 < -fd.nb[0] || s0 > fd.nb[0] || s1 < -fd.nb[1] || s1 > fd.nb[1] ||
            s2 < -fd.nb[2] || s2 > fd.nb[2])
          continue;  // image outside the reference's shift list (cannot hold a neighbour)
        const int img = img_pos[fd.img_off + ((s0 + fd.nb[0]) * w1 + (s1 + fd.nb[1])) * w2 +
                                (s2 + fd.nb[2])];
        const double* sv = shift_vec + 3 * (size_t)(fd.img_off + img);
        const double svx = sv[0], svy = sv[1], svz = sv[2];
        const bool zero_shift = (s0 == 0 && s1 == 0 && s2 == 0);
        const int cid = fd.cell_off + (u0 * fd.nc[1] + u1) * fd.nc[2] + u2;
        const int kb = cell_start[cid], ke = cell_start[cid + 1];
        for (int k0 = kb; k0 < ke; k0 += kWave) {
          const int k = k0 + lane;
          bool keep = false;
          double r = 0.0;
          int e = 0, tj = -1;
          if (k < ke) {
            const int j = cell_atoms[k];
            tj = atype[j];
            if (tj >= 0 && !(zero_shift && j == g)) {
              // extended coordinate = wrapped + shift vector (nlist.py:503)
              const double ex = __dadd_rn(wrapped[3 * j], svx);
              const double ey = __dadd_rn(wrapped[3 * j + 1], svy);
              const double ez = __dadd_rn(wrapped[3 * j + 2], svz);
              r = norm3_ref(__dsub_rn(ex, cx), __dsub_rn(ey, cy), __dsub_rn(ez, cz));
              keep = r <= rcut;
              e = img * fd.nloc + (j - fd.atom_off);
            }
          }
          const unsigned long long m = __ballot(keep);
          const int pos = n + lanes_below(m);
          if (keep && pos < MAXC) {
            cand[pos].r = r;
            cand[pos].e = e;
            cand[pos].t = tj;
          }
          n += __popcll(m);
        }
      }
    }
  }
  if (lane == 0) atomicMax(status, n);
  const int nc = n < MAXC ? n : MAXC;
  for (int s = lane; s < S; s += kWave) row[s] = -1;
  __syncthreads();
  // rank by (distance, extended index); keep the nnei nearest, then sel[t]
  for (int c = lane; c < nc; c += kWave) {
    const double r = cand[c].r;
    const int e = cand[c].e, t = cand[c].t;
    int rank = 0, trank = 0;
    for (int q = 0; q < nc; ++q) {
      const Cand o = cand[q];
      const bool less = (o.r < r) || (o.r == r && o.e < e);
      rank += less;
      trank += (less && o.t == t);
    }
    if (rank < S && t < si.ntypes && trank < si.sel[t]) row[si.sec[t] + trank] = e;
  }
  __syncthreads();
  for (int s = lane; s < S; s += kWave) out[s] = row[s];
}

template <int MAXC>
int launch_search(const double* wrapped, const int* atype, const int* frame_of,
                  const dpk_frame* frames, const int* atom_cell, const int* cell_start,
                  const int* cell_atoms, const int* img_pos, const double* shift_vec,
                  double rcut, const SelInfo& si, int natoms, int* nlist, int* status,
                  hipStream_t st) {
  size_t lds = sizeof(Cand) * MAXC + sizeof(int) * (size_t)si.nnei;
  DPK_REQUIRE(lds <= 64 * 1024, DPK_ERR_UNSUPPORTED, "nlist: LDS %zu bytes too large", lds);
  hipLaunchKernelGGL(k_search<MAXC>, dim3(natoms), dim3(64), lds, st, wrapped, atype,
                     frame_of, frames, atom_cell, cell_start, cell_atoms, img_pos, shift_vec,
                     rcut, si, natoms, nlist, status);
  DPK_LAUNCHED();
  return DPK_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace
}  // namespace dpk

using namespace dpk;

extern "C" size_t dpk_nlist_workspace_bytes(int natoms, int ncells) {
  return align256(sizeof(int) * (size_t)ncells) + align256(sizeof(int) * ((size_t)ncells + 1)) +
         3 * align256(sizeof(int) * (size_t)natoms);
}

extern "C" int dpk_nlist_build(int* nlist, double* coord_wrapped, int* status_dev,
                               const double* coord, const int* atype, const int* frame_of,
                               const dpk_frame* frames, int nframes, int natoms, int ncells,
                               const int* img_pos, const double* shift_vec, double rcut,
                               const int* sel_host, int ntypes, int max_cand, void* workspace,
                               size_t workspace_bytes, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && ncells > 0 && nframes > 0, DPK_ERR_ARG,
              "dpk_nlist_build: bad sizes natoms=%d ncells=%d nframes=%d", natoms, ncells, nframes);
  DPK_REQUIRE(ntypes > 0 && ntypes <= kMaxTypes, DPK_ERR_ARG, "dpk_nlist_build: ntypes=%d", ntypes);
  DPK_REQUIRE(workspace_bytes >= dpk_nlist_workspace_bytes(natoms, ncells), DPK_ERR_ARG,
              "dpk_nlist_build: workspace too small");
  if (natoms == 0) return DPK_OK;
  SelInfo si;
  si.ntypes = ntypes;
  si.sec[0] = 0;
  for (int t = 0; t < ntypes; ++t) {
    DPK_REQUIRE(sel_host[t] >= 0, DPK_ERR_ARG, "dpk_nlist_build: negative sel");
    si.sel[t] = sel_host[t];
    si.sec[t + 1] = si.sec[t] + sel_host[t];
  }
  si.nnei = si.sec[ntypes];
  DPK_REQUIRE(si.nnei > 0, DPK_ERR_ARG, "dpk_nlist_build: sum(sel) == 0");
  hipStream_t st = as_stream(stream);
  char* ws = static_cast<char*>(workspace);
  int* cell_count = reinterpret_cast<int*>(ws);
  ws += align256(sizeof(int) * (size_t)ncells);
  int* cell_start = reinterpret_cast<int*>(ws);
  ws += align256(sizeof(int) * ((size_t)ncells + 1));
  int* cell_atoms = reinterpret_cast<int*>(ws);
  ws += align256(sizeof(int) * (size_t)natoms);
  int* atom_cell = reinterpret_cast<int*>(ws);
  ws += align256(sizeof(int) * (size_t)natoms);
  int* atom_slot = reinterpret_cast<int*>(ws);

  hipLaunchKernelGGL(k_zero, dim3(cdiv(ncells, 256)), dim3(256), 0, st, cell_count, ncells,
                     status_dev);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_bin, dim3(cdiv(natoms, 256)), dim3(256), 0, st, coord, frame_of, frames,
                     natoms, coord_wrapped, atom_cell, atom_slot, cell_count);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cell_count, ncells, cell_start);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_scatter, dim3(cdiv(natoms, 256)), dim3(256), 0, st, atom_cell, atom_slot,
                     cell_start, natoms, cell_atoms);
  DPK_LAUNCHED();
#define DPK_SEARCH(M)                                                                       \
  return launch_search<M>(coord_wrapped, atype, frame_of, frames, atom_cell, cell_start,   \
                          cell_atoms, img_pos, shift_vec, rcut, si, natoms, nlist,         \
                          status_dev, st)
  if (max_cand <= 256) DPK_SEARCH(256);
  if (max_cand <= 512) DPK_SEARCH(512);
  if (max_cand <= 1024) DPK_SEARCH(1024);
  if (max_cand <= 2048) DPK_SEARCH(2048);
#undef DPK_SEARCH
  set_error("dpk_nlist_build: max_cand=%d above 2048", max_cand);
  return DPK_ERR_UNSUPPORTED;
}
