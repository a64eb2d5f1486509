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
#include <mutex>

#include "dpk_common.h"
#include "nlist_ws.h"

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
                          int* __restrict__ cell_tmp) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= natoms) return;
  cell_tmp[cell_start[atom_cell[g]] + atom_slot[g]] = g;
}

// Arrival order inside a cell (atomics) is not reproducible; each atom's final
// slot is its rank by index among the cell's atoms, so cell_atoms lists every
// cell in ascending atom index and every later kernel sees a fixed order.
__global__ void k_rank(const int* __restrict__ atom_cell, const int* __restrict__ cell_start,
                       const int* __restrict__ cell_tmp, int natoms, int* __restrict__ atom_slot,
                       int* __restrict__ cell_atoms) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= natoms) return;
  const int c = atom_cell[g];
  const int a = cell_start[c], b = cell_start[c + 1];
  int rank = 0;
  for (int k = a; k < b; ++k) rank += cell_tmp[k] < g ? 1 : 0;
  atom_slot[g] = rank;
  cell_atoms[a + rank] = g;
}

// r <= rcut for r = sqrt(r2) in FP64 (torch's rounding), with the square root
// taken only in the thin band where r2 alone cannot decide: lo = rcut^2 (1 -
// 1e-12) and hi = rcut^2 (1 + 1e-12) bracket every rounding of rcut^2 and of
// the square root.
__device__ __forceinline__ bool within_rcut(double r2, double lo, double hi, double rcut) {
  return r2 <= lo || (r2 <= hi && __dsqrt_rn(r2) <= rcut);
}

__device__ __forceinline__ int floordiv(int a, int b) {
  int q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ballot compaction of one 64-candidate chunk into the wavefront's buffer.
template <int MAXC>
__device__ __forceinline__ void push(Cand* cand, int& n, bool keep, double r, int e, int t,
                                     int lane) {
  const unsigned long long m = __ballot(keep);
  const int pos = n + lanes_below(m);
  if (keep && pos < MAXC) {
    cand[pos].r = r;
    cand[pos].e = e;
    cand[pos].t = t;
  }
  n += __popcll(m);
}

// Candidates of centre g from global memory (one wavefront):
//  1. the (2 rng + 1)^3 surrounding cells, 64 at a time, one per lane (cell
//     range, image, shift vector), with an inclusive prefix of their sizes;
//  2. the cells' atoms as one flat candidate index: lane q takes candidate
//     base + q (owner cell found by a 6-step binary search over the prefix),
//     tests r^2 against rcut^2 (with a relative margin, so every r <= rcut
//     passes) and only then takes the FP64 norm in torch's rounding and the
//     exact r <= rcut test; survivors are compacted into LDS by ballot.
template <int MAXC>
__device__ int gather_global(Cand* cand, const double* __restrict__ wrapped,
                             const int* __restrict__ atype, const dpk_frame& fd, int g,
                             const int* __restrict__ atom_cell, const int* __restrict__ cell_start,
                             const int* __restrict__ cell_atoms, const int* __restrict__ img_pos,
                             const double* __restrict__ shift_vec, double rcut, int lane) {
  const double cx = wrapped[3 * g], cy = wrapped[3 * g + 1], cz = wrapped[3 * g + 2];
  const int loc = atom_cell[g] - fd.cell_off;
  const int c2 = loc % fd.nc[2];
  const int c1 = (loc / fd.nc[2]) % fd.nc[1];
  const int c0 = loc / (fd.nc[2] * fd.nc[1]);
  const int w1 = 2 * fd.nb[1] + 1, w2 = 2 * fd.nb[2] + 1;
  const int span0 = 2 * fd.rng[0] + 1, span1 = 2 * fd.rng[1] + 1, span2 = 2 * fd.rng[2] + 1;
  const int ncell = span0 * span1 * span2;
  const double rc2 = rcut * rcut;
  const double lo = rc2 * (1.0 - 1e-12), hi = rc2 * (1.0 + 1e-12);
  int n = 0;
  for (int cb = 0; cb < ncell; cb += kWave) {
    // ---- one surrounding cell per lane
    int kb = 0, cnt = 0, img = 0;
    double svx = 0.0, svy = 0.0, svz = 0.0;
    bool zs = false;
    const int ci = cb + lane;
    if (ci < ncell) {
      const int d2 = ci % span2 - fd.rng[2];
      const int d1 = (ci / span2) % span1 - fd.rng[1];
      const int d0 = ci / (span2 * span1) - fd.rng[0];
      int u0 = c0 + d0, u1 = c1 + d1, u2 = c2 + d2, s0 = 0, s1 = 0, s2 = 0;
      bool ok = true;
      if (fd.pbc) {
        s0 = floordiv(u0, fd.nc[0]); u0 -= s0 * fd.nc[0];
        s1 = floordiv(u1, fd.nc[1]); u1 -= s1 * fd.nc[1];
        s2 = floordiv(u2, fd.nc[2]); u2 -= s2 * fd.nc[2];
        // images outside the reference's shift list cannot hold a neighbour
        ok = s0 >= -fd.nb[0] && s0 <= fd.nb[0] && s1 >= -fd.nb[1] && s1 <= fd.nb[1] &&
             s2 >= -fd.nb[2] && s2 <= fd.nb[2];
      } else {
        ok = u0 >= 0 && u0 < fd.nc[0] && u1 >= 0 && u1 < fd.nc[1] && u2 >= 0 && u2 < fd.nc[2];
      }
      if (ok) {
        img = img_pos[fd.img_off + ((s0 + fd.nb[0]) * w1 + (s1 + fd.nb[1])) * w2 +
                      (s2 + fd.nb[2])];
        const double* sv = shift_vec + 3 * (size_t)(fd.img_off + img);
        svx = sv[0];
        svy = sv[1];
        svz = sv[2];
        zs = (s0 == 0 && s1 == 0 && s2 == 0);
        const int cid = fd.cell_off + (u0 * fd.nc[1] + u1) * fd.nc[2] + u2;
        kb = cell_start[cid];
        cnt = cell_start[cid + 1] - kb;
      }
    }
    int pre = cnt;  // inclusive prefix over lanes
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int v = __shfl_up(pre, off, kWave);
      if (lane >= off) pre += v;
    }
    const int total = __shfl(pre, kWave - 1, kWave);
    // ---- flat candidate index over this group of cells
    for (int b0 = 0; b0 < total; b0 += kWave) {
      const int idx = b0 + lane;
      const int q = idx < total ? idx : total - 1;
      int own = 0;
#pragma unroll
      for (int step = kWave / 2; step > 0; step >>= 1)
        if (__shfl(pre, own + step - 1, kWave) <= q) own += step;
      const int first = __shfl(pre, own, kWave) - __shfl(cnt, own, kWave);
      const int k = __shfl(kb, own, kWave) + (q - first);
      const int im = __shfl(img, own, kWave);
      const double ox = __shfl(svx, own, kWave), oy = __shfl(svy, own, kWave),
                   oz = __shfl(svz, own, kWave);
      const bool ozs = __shfl((int)zs, own, kWave) != 0;
      bool keep = false;
      double r = 0.0;
      int e = 0, tj = -1;
      if (idx < total) {
        const int j = cell_atoms[k];
        tj = atype[j];
        if (tj >= 0 && !(ozs && j == g)) {
          // extended coordinate = wrapped + shift vector (nlist.py:503)
          const double dx = __dsub_rn(__dadd_rn(wrapped[3 * j], ox), cx);
          const double dy = __dsub_rn(__dadd_rn(wrapped[3 * j + 1], oy), cy);
          const double dz = __dsub_rn(__dadd_rn(wrapped[3 * j + 2], oz), cz);
          r = __fma_rn(dz, dz, __fma_rn(dy, dy, __dmul_rn(dx, dx)));  // r^2, the key's square
          keep = within_rcut(r, lo, hi, rcut);
          e = im * fd.nloc + (j - fd.atom_off);
        }
      }
      push<MAXC>(cand, n, keep, r, e, tj, lane);
    }
  }
  return n;
}

// Step 4 of both sorts below: in sorted order, the first sum(sel) of the nc
// candidates are dealt to their type blocks (ballot prefix counts per type),
// each truncated at sel[t]; the row is assembled in LDS and written out.
__device__ void deal_and_write(const Cand* cand, int nc, int* row, int* tbase, const SelInfo& si,
                               int* __restrict__ out, int lane) {
  const int S = si.nnei;
  for (int s = lane; s < S; s += kWave) row[s] = -1;
  if (lane < si.ntypes) tbase[lane] = 0;
  wave_sync();
  const int lim = nc < S ? nc : S;
  for (int s0 = 0; s0 < lim; s0 += kWave) {
    const int pos = s0 + lane;
    const bool valid = pos < lim;
    const Cand c = cand[valid ? pos : 0];
    for (int t = 0; t < si.ntypes; ++t) {
      const bool mine = valid && c.t == t;
      const unsigned long long m = __ballot(mine);
      if (m == 0) continue;
      const int base = tbase[t];
      if (mine) {
        const int trank = base + lanes_below(m);
        if (trank < si.sel[t]) row[si.sec[t] + trank] = c.e;
      }
      wave_sync();
      if (lane == 0) tbase[t] = base + __popcll(m);
      wave_sync();
    }
  }
  wave_sync();
  for (int s = lane; s < S; s += kWave) out[s] = row[s];
}

// Candidates [0, n) of one wavefront (key field = r^2) sorted in LDS and
// written as the nlist row: 3. bitonic sort by (r, extended index), r =
// sqrt(r^2) in FP64 (skipped when do_sort is false: the row keeps the
// candidates' order); 4. deal_and_write.
template <int MAXC>
__device__ void sort_and_write(Cand* cand, int n, int* row, int* tbase, const SelInfo& si,
                               int* __restrict__ out, int lane, bool do_sort = true) {
  const int nc = n < MAXC ? n : MAXC;
  int npad = kWave;
  while (npad < nc) npad <<= 1;
  for (int c = nc + lane; c < npad; c += kWave) {
    cand[c].r = __longlong_as_double(0x7ff0000000000000LL);  // +inf: padding sorts last
    cand[c].e = 0x7fffffff;
    cand[c].t = -1;
  }
  if (do_sort)  // candidates carry r^2; the key is the FP64 distance itself
    for (int c = lane; c < nc; c += kWave) cand[c].r = __dsqrt_rn(cand[c].r);
  wave_sync();
  // every lane takes compare pairs (i, i + jj): pair index w -> i by inserting
  // a zero bit at jj, so no lane idles
  for (int kk = 2; do_sort && kk <= npad; kk <<= 1) {
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      for (int w = lane; w < npad / 2; w += kWave) {
        const int i = ((w & ~(jj - 1)) << 1) | (w & (jj - 1));
        const int p = i | jj;
        const Cand a = cand[i], b = cand[p];
        const bool agtb = (a.r > b.r) || (a.r == b.r && a.e > b.e);
        if (agtb == ((i & kk) == 0)) {
          cand[i] = b;
          cand[p] = a;
        }
      }
      wave_sync();
    }
  }
  deal_and_write(cand, nc, row, tbase, si, out, lane);
}

// The in-register stages of bitonic block kk: distances J = EPL/2 .. 1 below
// kk, between positions base + k and base + (k | J) of one lane, ascending
// where (pos & kk) == 0.  J and k are compile-time constants, so the arrays
// stay in registers (a runtime distance would index them dynamically).
template <int EPL>
__device__ __forceinline__ void reg_stages(double (&key)[EPL], int (&pay)[EPL], int base, int kk) {
#pragma unroll
  for (int J = EPL / 2; J >= 1; J >>= 1) {
    if (J < kk) {
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        if ((k & J) == 0) {
          const int p = k | J;
          const bool asc = ((base + k) & kk) == 0;
          const bool gt = key[k] > key[p] || (key[k] == key[p] && pay[k] > pay[p]);
          if (gt == asc) {
            const double tk = key[k];
            const int tp = pay[k];
            key[k] = key[p];
            pay[k] = pay[p];
            key[p] = tk;
            pay[p] = tp;
          }
        }
      }
    }
  }
}

// Candidates [0, n) (key field = r^2) sorted in registers and written as the
// nlist row.  Lane l holds positions EPL l .. EPL l + EPL - 1 (key = FP64
// distance, payload = extended index); compare partners closer than EPL are
// in the same lane, farther ones are exchanged with lane shuffles -- no LDS
// traffic in the network.  Types are looked up again for the dealing step.
template <int MAXC>
__device__ void sort_regs_and_write(Cand* cand, int n, int* row, int* tbase, const SelInfo& si,
                                    int* __restrict__ out, const int* __restrict__ atype,
                                    int atom_off, int nloc, int lane) {
  constexpr int EPL = MAXC / kWave;
  const int nc = n < MAXC ? n : MAXC;
  int npad = EPL;
  while (npad < nc) npad <<= 1;
  const int base = EPL * lane;
  double key[EPL];
  int pay[EPL];
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    const bool real = base + k < nc;
    key[k] = real ? __dsqrt_rn(cand[base + k].r) : __longlong_as_double(0x7ff0000000000000LL);
    pay[k] = real ? cand[base + k].e : 0x7fffffff;
  }
  for (int kk = 2; kk <= npad; kk <<= 1) {
    for (int jj = kk >> 1; jj >= EPL; jj >>= 1) {
      const int lm = jj / EPL;
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const double q = __shfl_xor(key[k], lm, kWave);
        const int qp = __shfl_xor(pay[k], lm, kWave);
        const bool lower = ((base + k) & jj) == 0;
        const bool asc = ((base + k) & kk) == 0;
        const bool gt = key[k] > q || (key[k] == q && pay[k] > qp);
        const bool lt = key[k] < q || (key[k] == q && pay[k] < qp);
        if (lower == asc ? gt : lt) {
          key[k] = q;
          pay[k] = qp;
        }
      }
    }
    reg_stages<EPL>(key, pay, base, kk);
  }
  wave_sync();  // every lane has read its candidates before they are overwritten
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    if (base + k < nc) {
      const int e = pay[k];
      const int im = e / nloc;
      cand[base + k].r = key[k];
      cand[base + k].e = e;
      cand[base + k].t = atype[atom_off + (e - im * nloc)];
    }
  }
  wave_sync();
  deal_and_write(cand, nc, row, tbase, si, out, lane);
}

// reg_stages on packed keys.
template <int EPL>
__device__ __forceinline__ void preg_stages(unsigned long long (&key)[EPL], int base, int kk) {
#pragma unroll
  for (int J = EPL / 2; J >= 1; J >>= 1) {
    if (J < kk) {
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        if ((k & J) == 0) {
          const int p = k | J;
          const bool asc = ((base + k) & kk) == 0;
          const unsigned long long a = key[k], b = key[p];
          const bool swap = (a > b) == asc;
          key[k] = swap ? b : a;
          key[p] = swap ? a : b;
        }
      }
    }
  }
}

// Bits of the extended index of a frame: e < ns * nloc.
__device__ __forceinline__ int index_bits(const dpk_frame& fd) {
  const unsigned long long emax = (unsigned long long)fd.ns * (unsigned long long)fd.nloc;
  return emax > 1 ? 64 - __clzll(emax - 1) : 1;
}

// The register sort on one 64-bit key per candidate: the bit pattern of the
// FP64 distance with its low `eb` bits replaced by the extended index
// (< 2^eb).  Non-negative doubles order like their bit patterns, so when no
// two neighbours of the sorted order share the distance part (the bits above
// eb), that order is exactly the (distance, index) order of the full-key
// sort.  Otherwise (equal distances, or distances equal to ~2^(eb-52)
// relative) it returns false before writing anything, and the caller runs
// the full-key sort on the untouched candidates.  A key is one 64-bit
// compare and two lane shuffles per exchange instead of an FP64 compare, an
// index compare and three shuffles.
template <int MAXC>
__device__ bool sort_packed_and_write(Cand* cand, int n, int* row, int* tbase, const SelInfo& si,
                                      int* __restrict__ out, const int* __restrict__ atype,
                                      int atom_off, const FastDiv& nloc_div, int lane, int eb) {
  constexpr int EPL = MAXC / kWave;
  const int nc = n < MAXC ? n : MAXC;
  int npad = EPL;
  while (npad < nc) npad <<= 1;
  const int base = EPL * lane;
  const unsigned long long emask = (1ull << eb) - 1;
  unsigned long long key[EPL];
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    const bool real = base + k < nc;
    key[k] = real ? (((unsigned long long)__double_as_longlong(__dsqrt_rn(cand[base + k].r)) &
                      ~emask) |
                     (unsigned)cand[base + k].e)
                  : ~0ull;  // padding sorts last (real keys have the sign bit clear)
  }
  for (int kk = 2; kk <= npad; kk <<= 1) {
    for (int jj = kk >> 1; jj >= EPL; jj >>= 1) {
      // partner lane l ^ (jj / EPL); every position of a lane has the same
      // direction bits (jj, kk >= EPL)
      const int lm = jj / EPL;
      const bool keep_min = ((base & jj) == 0) == ((base & kk) == 0);
#pragma unroll
      for (int k = 0; k < EPL; ++k) {
        const unsigned long long q = __shfl_xor(key[k], lm, kWave);
        if ((key[k] > q) == keep_min) key[k] = q;
      }
    }
    preg_stages<EPL>(key, base, kk);
  }
  // neighbours in sorted order with the same distance part?
  bool tie = false;
#pragma unroll
  for (int k = 0; k + 1 < EPL; ++k)
    tie |= base + k + 1 < nc && (key[k] >> eb) == (key[k + 1] >> eb);
  const unsigned long long nx = __shfl_down(key[0], 1, kWave);
  tie |= lane + 1 < kWave && base + EPL < nc && (key[EPL - 1] >> eb) == (nx >> eb);
  if (__ballot(tie) != 0) return false;
  wave_sync();  // every lane has read its candidates before they are overwritten
#pragma unroll
  for (int k = 0; k < EPL; ++k) {
    if (base + k < nc) {
      const unsigned e = (unsigned)(key[k] & emask);
      cand[base + k].e = (int)e;
      cand[base + k].t = atype[atom_off + (int)(e - nloc_div.div(e) * nloc_div.d)];
    }
  }
  wave_sync();
  deal_and_write(cand, nc, row, tbase, si, out, lane);
  return true;
}

// Block-staged candidate of a cell's 27-cell neighbourhood (extended position)
struct BCand {
  double x, y, z;
  int e, t;
};
constexpr int kBlockCells = 27;

template <int MAXC>
__host__ __device__ constexpr size_t search_wave_bytes(int nnei) {
  return ((sizeof(Cand) * MAXC + sizeof(int) * (size_t)(nnei + kMaxTypes)) + 15) / 16 * 16;
}

// One workgroup per cell of the grid; its wavefronts take the cell's atoms as
// centres.  The candidates of the 27 surrounding cells (extended position,
// extended index, type) are staged in LDS once for the whole cell, so a
// centre scans LDS instead of gathering from global memory.  Cells whose
// neighbourhood does not fit (more than `capb` candidates, or rng > 1 in
// small boxes) use the per-centre global gather.
template <int MAXC, int W>
__global__ void __launch_bounds__(64 * W) k_search_cells(
    const double* __restrict__ wrapped, const int* __restrict__ atype,
    const int* __restrict__ frame_of, const dpk_frame* __restrict__ frames,
    const int* __restrict__ atom_cell, const int* __restrict__ cell_start,
    const int* __restrict__ cell_atoms, const int* __restrict__ img_pos,
    const double* __restrict__ shift_vec, double rcut, SelInfo si, int capb, bool packed,
    int* __restrict__ nlist, int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int c_kb[kBlockCells], c_pre[kBlockCells + 1], c_img[kBlockCells];
  __shared__ int wg_max;
  const int cell = blockIdx.x;
  const int a0 = cell_start[cell], a1 = cell_start[cell + 1];
  if (a1 <= a0) return;  // uniform over the workgroup
  if (threadIdx.x == 0) wg_max = 0;
  __syncthreads();
  int wmax = 0;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int S = si.nnei;
  BCand* bc = reinterpret_cast<BCand*>(smem);
  unsigned char* wb = smem + sizeof(BCand) * (size_t)capb + search_wave_bytes<MAXC>(S) * wave;
  Cand* cand = reinterpret_cast<Cand*>(wb);
  int* row = reinterpret_cast<int*>(wb + sizeof(Cand) * MAXC);
  int* tbase = row + S;
  const dpk_frame& fd = frames[frame_of[cell_atoms[a0]]];
  const int loc = cell - fd.cell_off;
  const int c2 = loc % fd.nc[2], c1 = (loc / fd.nc[2]) % fd.nc[1],
            c0 = loc / (fd.nc[2] * fd.nc[1]);
  const bool small = fd.rng[0] > 1 || fd.rng[1] > 1 || fd.rng[2] > 1;
  const int w1 = 2 * fd.nb[1] + 1, w2 = 2 * fd.nb[2] + 1;
  if (!small) {
    for (int k = threadIdx.x; k < kBlockCells; k += blockDim.x) {
      const int d0 = k / 9 - 1, d1 = (k / 3) % 3 - 1, d2 = k % 3 - 1;
      int u0 = c0 + d0, u1 = c1 + d1, u2 = c2 + d2, s0 = 0, s1 = 0, s2 = 0;
      bool ok = true;
      if (fd.pbc) {
        s0 = floordiv(u0, fd.nc[0]); u0 -= s0 * fd.nc[0];
        s1 = floordiv(u1, fd.nc[1]); u1 -= s1 * fd.nc[1];
        s2 = floordiv(u2, fd.nc[2]); u2 -= s2 * fd.nc[2];
        ok = s0 >= -fd.nb[0] && s0 <= fd.nb[0] && s1 >= -fd.nb[1] && s1 <= fd.nb[1] &&
             s2 >= -fd.nb[2] && s2 <= fd.nb[2];
      } else {
        ok = u0 >= 0 && u0 < fd.nc[0] && u1 >= 0 && u1 < fd.nc[1] && u2 >= 0 && u2 < fd.nc[2];
      }
      int kb = 0, cnt = 0, img = 0;
      if (ok) {
        img = img_pos[fd.img_off + ((s0 + fd.nb[0]) * w1 + (s1 + fd.nb[1])) * w2 +
                      (s2 + fd.nb[2])];
        const int cid = fd.cell_off + (u0 * fd.nc[1] + u1) * fd.nc[2] + u2;
        kb = cell_start[cid];
        cnt = cell_start[cid + 1] - kb;
      }
      c_kb[k] = kb;
      c_img[k] = img;
      c_pre[k + 1] = cnt;  // turned into a prefix below
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      c_pre[0] = 0;
      for (int k = 0; k < kBlockCells; ++k) c_pre[k + 1] += c_pre[k];
    }
    __syncthreads();
  }
  const int total = small ? capb + 1 : c_pre[kBlockCells];
  const bool staged = total <= capb;
  if (staged) {
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
      int own = 0;  // last cell with c_pre[own] <= q
#pragma unroll
      for (int step = 16; step > 0; step >>= 1)
        if (own + step <= kBlockCells - 1 && c_pre[own + step] <= q) own += step;
      while (c_pre[own + 1] <= q) ++own;  // skip empty cells
      const int j = cell_atoms[c_kb[own] + (q - c_pre[own])];
      const int im = c_img[own];
      const double* sv = shift_vec + 3 * (size_t)(fd.img_off + im);
      BCand b;
      b.x = __dadd_rn(wrapped[3 * j], sv[0]);
      b.y = __dadd_rn(wrapped[3 * j + 1], sv[1]);
      b.z = __dadd_rn(wrapped[3 * j + 2], sv[2]);
      b.e = im * fd.nloc + (j - fd.atom_off);
      b.t = atype[j];
      bc[q] = b;
    }
    __syncthreads();
  }
  const double rc2 = rcut * rcut;
  const double lo = rc2 * (1.0 - 1e-12), hi = rc2 * (1.0 + 1e-12);
  const int eb = index_bits(fd);
  FastDiv nloc_div;
  nloc_div.init((unsigned)fd.nloc);
  for (int ai = a0 + wave; ai < a1; ai += W) {
    const int g = cell_atoms[ai];
    int* out = nlist + (size_t)g * S;
    if (atype[g] < 0) {  // virtual centre: empty row (nlist.py:193-195)
      for (int s = lane; s < S; s += kWave) out[s] = -1;
      continue;
    }
    int n = 0;
    if (staged) {
      const double cx = wrapped[3 * g], cy = wrapped[3 * g + 1], cz = wrapped[3 * g + 2];
      const int eself = g - fd.atom_off;  // image 0 is the zero shift
      for (int q0 = 0; q0 < total; q0 += kWave) {
        const int q = q0 + lane;
        bool keep = false;
        double r = 0.0;
        int e = 0, t = -1;
        if (q < total) {
          const BCand b = bc[q];
          if (b.t >= 0 && b.e != eself) {
            const double dx = __dsub_rn(b.x, cx);
            const double dy = __dsub_rn(b.y, cy);
            const double dz = __dsub_rn(b.z, cz);
            r = __fma_rn(dz, dz, __fma_rn(dy, dy, __dmul_rn(dx, dx)));
            keep = within_rcut(r, lo, hi, rcut);
            e = b.e;
            t = b.t;
          }
        }
        push<MAXC>(cand, n, keep, r, e, t, lane);
      }
    } else {
      n = gather_global<MAXC>(cand, wrapped, atype, fd, g, atom_cell, cell_start, cell_atoms,
                              img_pos, shift_vec, rcut, lane);
    }
    // register sort up to 8 candidates per lane; larger capacities sort in LDS
    if constexpr (MAXC <= 512) {
      if (!(packed && sort_packed_and_write<MAXC>(cand, n, row, tbase, si, out, atype,
                                                  fd.atom_off, nloc_div, lane, eb)))
        sort_regs_and_write<MAXC>(cand, n, row, tbase, si, out, atype, fd.atom_off, fd.nloc, lane);
    } else
      sort_and_write<MAXC>(cand, n, row, tbase, si, out, lane);
    wmax = n > wmax ? n : wmax;
  }
  // largest candidate count: one global atomic per workgroup (a same-address
  // atomic per centre serialises in L2)
  if (lane == 0) atomicMax(&wg_max, wmax);
  __syncthreads();
  if (threadIdx.x == 0) atomicMax(status, wg_max);
}

template <int MAXC, int W>
int launch_search_w(const double* wrapped, const int* atype, const int* frame_of,
                    const dpk_frame* frames, const int* atom_cell, const int* cell_start,
                    const int* cell_atoms, const int* img_pos, const double* shift_vec,
                    double rcut, const SelInfo& si, int natoms, int ncells, int* nlist,
                    int* status, size_t budget, hipStream_t st) {
  // staged candidates: ~1.25x the mean 27-cell neighbourhood, within the LDS budget
  const size_t wbytes = (size_t)W * search_wave_bytes<MAXC>(si.nnei);
  DPK_REQUIRE(wbytes + 64 * sizeof(BCand) <= budget, DPK_ERR_UNSUPPORTED,
              "nlist: candidate buffers of %zu bytes exceed the LDS budget %zu", wbytes, budget);
  long want = (long)(27.0 * natoms / ncells * tune(DPK_TUNE_NLIST_STAGE_PCT) / 100.0) + 64;
  long capb = (long)((budget - wbytes) / sizeof(BCand));
  if (want < capb) capb = want;
  capb = capb / 64 * 64;
  if (capb < 64) capb = 64;
  if (tune(DPK_TUNE_NLIST_MODE) == 1) capb = 0;  // every centre gathers from global memory
  const size_t lds = sizeof(BCand) * (size_t)(capb > 0 ? capb : 1) + wbytes;
  hipLaunchKernelGGL((k_search_cells<MAXC, W>), dim3(ncells), dim3(64 * W), lds, st, wrapped,
                     atype, frame_of, frames, atom_cell, cell_start, cell_atoms, img_pos,
                     shift_vec, rcut, si, (int)capb, tune(DPK_TUNE_NLIST_SORT) == 0, nlist,
                     status);
  DPK_LAUNCHED();
  return DPK_OK;
}

// Wavefronts per cell workgroup: more of them share one staged neighbourhood
// and hide the latency of the per-centre scan and sort (a workgroup of 16
// takes a whole CU); fewer where cells hold few atoms.
template <int MAXC>
int launch_search(const double* wrapped, const int* atype, const int* frame_of,
                  const dpk_frame* frames, const int* atom_cell, const int* cell_start,
                  const int* cell_atoms, const int* img_pos, const double* shift_vec,
                  double rcut, const SelInfo& si, int natoms, int ncells, int* nlist,
                  int* status, hipStream_t st) {
  int dev = 0, maxlds = 0;
  DPK_HIP(hipGetDevice(&dev));
  DPK_HIP(hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  const size_t budget = (size_t)(maxlds < 160 * 1024 ? maxlds : 160 * 1024) - 512;
  const double per_cell = (double)natoms / ncells;
  int w = tune(DPK_TUNE_NLIST_WAVES);
  if (w == 0) w = per_cell >= 32.0 ? 16 : (per_cell >= 16.0 ? 8 : 4);
  w = w <= 4 ? 4 : (w <= 8 ? 8 : 16);
  // keep room for the per-wavefront buffers and a full mean neighbourhood
  const size_t stage = (size_t)(27.0 * per_cell + 64) * sizeof(BCand);
  while (w > 4 && (size_t)w * search_wave_bytes<MAXC>(si.nnei) + stage > budget) w /= 2;
#define DPK_SEARCH_W(NW)                                                                        \
  return launch_search_w<MAXC, NW>(wrapped, atype, frame_of, frames, atom_cell, cell_start,     \
                                   cell_atoms, img_pos, shift_vec, rcut, si, natoms, ncells,    \
                                   nlist, status, budget, st)
  if (w == 16) DPK_SEARCH_W(16);
  if (w == 8) DPK_SEARCH_W(8);
  DPK_SEARCH_W(4);
#undef DPK_SEARCH_W
}

// ------------------------------------------------- external nlist format --
// deepmd/pt/model/model/make_model.py:592-658 (_format_nlist) followed by
// nlist.py:304-337 (nlist_distinguish_types) on a caller's neighbour list
// (e.g. an MD engine's, built with rcut + skin, unsorted): one wavefront per
// centre row reads the row in order, drops entries beyond rcut (FP64 norm in
// torch's CPU rounding) and, when the row is wider than sum(sel) or a sort is
// requested, sorts the rest by (distance, index) and keeps the sum(sel)
// nearest; the survivors are then dealt to the type blocks in order and cut
// at sel[t].  Entries of type < 0 take part in the sort and the sum(sel) cut
// but belong to no block, as in the reference.

__global__ void k_clear(int* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <int MAXC, int W>
__global__ void __launch_bounds__(64 * W) k_format(
    const int* __restrict__ nlist_in, int nnei_in, const double* __restrict__ ext_coord,
    const int* __restrict__ ext_atype, int nrows, int nloc, int nall, double rcut, SelInfo si,
    int do_sort, int* __restrict__ nlist, int* __restrict__ status) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int wg_max;
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) wg_max = 0;
  __syncthreads();
  unsigned char* wb = smem + (size_t)wave * search_wave_bytes<MAXC>(si.nnei);
  Cand* cand = reinterpret_cast<Cand*>(wb);
  int* rbuf = reinterpret_cast<int*>(wb + sizeof(Cand) * MAXC);
  int* tbase = rbuf + si.nnei;
  const double lo = rcut * rcut * (1.0 - 1e-12), hi = rcut * rcut * (1.0 + 1e-12);
  int wmax = 0;
  // rows are strided over the grid (a few workgroups per CU), so the largest
  // count takes one global atomic per workgroup, not one per row
  for (int row = blockIdx.x * W + wave; row < nrows; row += gridDim.x * W) {
    const int f = row / nloc, i = row - f * nloc;
    const double* xf = ext_coord + (size_t)f * nall * 3;
    const int* tf = ext_atype + (size_t)f * nall;
    const double cx = xf[3 * i], cy = xf[3 * i + 1], cz = xf[3 * i + 2];
    const int* in = nlist_in + (size_t)row * nnei_in;
    int n = 0;
    // four chunks of 64 entries per lane at a time: the list reads, then the
    // coordinate and type gathers, are issued before the distances
    constexpr int kBatch = 4;
    for (int q0 = 0; q0 < nnei_in; q0 += kBatch * kWave) {
      int jv[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int q = q0 + u * kWave + lane;
        jv[u] = q < nnei_in ? in[q] : -1;
      }
      double px[kBatch], py[kBatch], pz[kBatch];
      int tv[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int j = jv[u];
        const bool ok = j >= 0 && j < nall;
        px[u] = ok ? xf[3 * j] : 0.0;
        py[u] = ok ? xf[3 * j + 1] : 0.0;
        pz[u] = ok ? xf[3 * j + 2] : 0.0;
        tv[u] = ok ? tf[j] : -1;
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int j = jv[u];
        bool keep = false;
        double r = 0.0;
        if (j >= nall) status[1] = 1;  // index outside the extended frame: reported, dropped
        if (j >= 0 && j < nall) {
          const double dx = __dsub_rn(cx, px[u]);
          const double dy = __dsub_rn(cy, py[u]);
          const double dz = __dsub_rn(cz, pz[u]);
          r = __fma_rn(dz, dz, __fma_rn(dy, dy, __dmul_rn(dx, dx)));
          keep = within_rcut(r, lo, hi, rcut);
        }
        // chunks past the row end push nothing (no lane keeps)
        push<MAXC>(cand, n, keep, r, keep ? j : 0, keep ? tv[u] : -1, lane);
      }
    }
    int* out = nlist + (size_t)row * si.nnei;
    // rows up to 512 wide sort in registers (types looked up again from the
    // extended index: atom_off 0, one image of nall atoms); wider ones in LDS
    if constexpr (MAXC <= 512) {
      if (do_sort)
        sort_regs_and_write<MAXC>(cand, n, rbuf, tbase, si, out, tf, 0, nall, lane);
      else
        sort_and_write<MAXC>(cand, n, rbuf, tbase, si, out, lane, false);
    } else {
      sort_and_write<MAXC>(cand, n, rbuf, tbase, si, out, lane, do_sort != 0);
    }
    wmax = n > wmax ? n : wmax;
  }
  if (lane == 0) atomicMax(&wg_max, wmax);
  __syncthreads();
  if (threadIdx.x == 0 && wg_max > 0) atomicMax(status, wg_max);
}

template <int MAXC, int W>
int launch_format(const int* nlist_in, int nnei_in, const double* ext_coord,
                  const int* ext_atype, int nrows, int nloc, int nall, double rcut,
                  const SelInfo& si, int do_sort, int* nlist, int* status, hipStream_t st) {
  const size_t lds = (size_t)W * search_wave_bytes<MAXC>(si.nnei);
  int dev = 0, ncu = 0, resident = 0;
  DPK_HIP(hipGetDevice(&dev));
  DPK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // resident workgroups per CU (registers and LDS): one round of the grid
  DPK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_format<MAXC, W>, 64 * W, lds));
  const long per_cu = resident > 0 ? resident : 1;
  long grid = cdiv(nrows, W);
  if (grid > per_cu * ncu) grid = per_cu * ncu;
  hipLaunchKernelGGL((k_format<MAXC, W>), dim3((unsigned)grid), dim3(64 * W), lds, st, nlist_in,
                     nnei_in, ext_coord, ext_atype, nrows, nloc, nall, rcut, si, do_sort, nlist,
                     status);
  DPK_LAUNCHED();
  return DPK_OK;
}

// Model-level pair exclusion: an entry whose (centre type, neighbour type)
// pair is excluded becomes -1 in place (a padding slot), after the list has
// been formatted.  A virtual centre (type < 0) uses row ntypes.
__global__ void k_pair_exclude(int* __restrict__ nlist, const int* __restrict__ atype,
                               const int* __restrict__ frame_of,
                               const dpk_frame* __restrict__ frames, int natoms, int nnei,
                               const int* __restrict__ pair_mask, int ntypes) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)natoms * nnei) return;
  const int e = nlist[q];
  if (e < 0) return;
  const int i = (int)(q / nnei);
  const dpk_frame& fr = frames[frame_of[i]];
  int ti = atype[i];
  ti = ti < 0 ? ntypes : ti;
  int tj = atype[fr.atom_off + e % fr.nloc];
  tj = tj < 0 ? ntypes : tj;
  if (pair_mask[ti * (ntypes + 1) + tj] == 0) nlist[q] = -1;
}

int sel_info(SelInfo& si, const int* sel_host, int ntypes, const char* who) {
  DPK_REQUIRE(ntypes > 0 && ntypes <= kMaxTypes, DPK_ERR_ARG, "%s: ntypes=%d", who, ntypes);
  si.ntypes = ntypes;
  si.sec[0] = 0;
  for (int t = 0; t < ntypes; ++t) {
    DPK_REQUIRE(sel_host[t] >= 0, DPK_ERR_ARG, "%s: negative sel", who);
    si.sel[t] = sel_host[t];
    si.sec[t + 1] = si.sec[t] + sel_host[t];
  }
  si.nnei = si.sec[ntypes];
  DPK_REQUIRE(si.nnei > 0, DPK_ERR_ARG, "%s: sum(sel) == 0", who);
  return DPK_OK;
}

}  // namespace
}  // namespace dpk

using namespace dpk;

extern "C" int dpk_nlist_format(int* nlist, int* status_dev, const int* nlist_in, int nnei_in,
                                const double* ext_coord, const int* ext_atype, int nframes,
                                int nloc, int nall, double rcut, const int* sel_host, int ntypes,
                                int extra_sort, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nframes >= 0 && nloc >= 0 && nall >= nloc && nnei_in >= 0, DPK_ERR_ARG,
              "dpk_nlist_format: bad sizes nframes=%d nloc=%d nall=%d nnei_in=%d", nframes,
              nloc, nall, nnei_in);
  SelInfo si;
  int rc = sel_info(si, sel_host, ntypes, "dpk_nlist_format");
  if (rc != DPK_OK) return rc;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_clear, dim3(1), dim3(64), 0, st, status_dev, 2);
  DPK_LAUNCHED();
  const int nrows = nframes * nloc;
  if (nrows == 0) return DPK_OK;
  const int do_sort = (extra_sort != 0 || nnei_in > si.nnei) ? 1 : 0;
#define DPK_FORMAT(M, W)                                                                     \
  return launch_format<M, W>(nlist_in, nnei_in, ext_coord, ext_atype, nrows, nloc, nall, rcut, \
                             si, do_sort, nlist, status_dev, st)
  if (nnei_in <= 256) DPK_FORMAT(256, 4);
  if (nnei_in <= 512) DPK_FORMAT(512, 4);
  if (nnei_in <= 1024) DPK_FORMAT(1024, 4);
  if (nnei_in <= 2048) DPK_FORMAT(2048, 2);
#undef DPK_FORMAT
  set_error("dpk_nlist_format: input rows of %d entries (at most 2048)", nnei_in);
  return DPK_ERR_UNSUPPORTED;
}

extern "C" size_t dpk_nlist_workspace_bytes(int natoms, int ncells) {
  return nlist_ws_bytes(natoms, ncells);
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
  DPK_REQUIRE(workspace_bytes >= dpk_nlist_workspace_bytes(natoms, ncells), DPK_ERR_ARG,
              "dpk_nlist_build: workspace too small");
  SelInfo si;
  const int rc = sel_info(si, sel_host, ntypes, "dpk_nlist_build");
  if (rc != DPK_OK) return rc;
  if (natoms == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  const NlistWs w = nlist_ws(workspace, natoms, ncells);
  int* cell_count = w.cell_count;
  int* cell_start = w.cell_start;
  int* cell_atoms = w.cell_atoms;
  int* atom_cell = w.atom_cell;
  int* atom_slot = w.atom_slot;

  hipLaunchKernelGGL(k_zero, dim3(cdiv(ncells, 256)), dim3(256), 0, st, cell_count, ncells,
                     status_dev);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_bin, dim3(cdiv(natoms, 256)), dim3(256), 0, st, coord, frame_of, frames,
                     natoms, coord_wrapped, atom_cell, atom_slot, cell_count);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, cell_count, ncells, cell_start);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_scatter, dim3(cdiv(natoms, 256)), dim3(256), 0, st, atom_cell, atom_slot,
                     cell_start, natoms, w.cell_tmp);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_rank, dim3(cdiv(natoms, 256)), dim3(256), 0, st, atom_cell, cell_start,
                     w.cell_tmp, natoms, atom_slot, cell_atoms);
  DPK_LAUNCHED();
#define DPK_SEARCH(M)                                                                       \
  return launch_search<M>(coord_wrapped, atype, frame_of, frames, atom_cell, cell_start,   \
                          cell_atoms, img_pos, shift_vec, rcut, si, natoms, ncells, nlist, \
                          status_dev, st)
  if (max_cand <= 256) DPK_SEARCH(256);
  if (max_cand <= 512) DPK_SEARCH(512);
  if (max_cand <= 1024) DPK_SEARCH(1024);
  if (max_cand <= 2048) DPK_SEARCH(2048);
#undef DPK_SEARCH
  set_error("dpk_nlist_build: max_cand=%d above 2048", max_cand);
  return DPK_ERR_UNSUPPORTED;
}

// ----------------------------------------------------- allocation-free --
// dpk_nlist_build_noalloc keeps its cell-list workspace in a file-scope
// device array (one copy per GPU, loaded with the code object and held until
// the process ends).  Host code never passes the array's name to a kernel:
// its device address is looked up once per device (hipGetSymbolAddress, with
// the device current) and that pointer is used.  Every call on a GPU shares
// the array, so calls that overlap on one GPU (two streams, two threads) race;
// each call rewrites every entry it reads (cell counts are cleared by k_zero).
constexpr size_t kNoallocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kNoallocMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kNoallocMaxDevices] = {};

int scratch_address(void** out) {
  int dev = 0;
  DPK_HIP(hipGetDevice(&dev));
  DPK_REQUIRE(dev >= 0 && dev < kNoallocMaxDevices, DPK_ERR_ARG,
              "dpk_nlist_build_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    DPK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    DPK_REQUIRE(p != nullptr, DPK_ERR_HIP, "dpk_nlist_build_noalloc: no scratch address");
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return DPK_OK;
}
}  // namespace

extern "C" int dpk_nlist_build_noalloc(int* nlist, double* coord_wrapped, int* status_dev,
                                       const double* coord, const int* atype,
                                       const int* frame_of, const dpk_frame* frames, int nframes,
                                       int natoms, int ncells, const int* img_pos,
                                       const double* shift_vec, double rcut,
                                       const int* sel_host, int ntypes, int max_cand,
                                       dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && ncells > 0 && nframes > 0, DPK_ERR_ARG,
              "dpk_nlist_build_noalloc: bad sizes natoms=%d ncells=%d nframes=%d", natoms, ncells,
              nframes);
  const size_t need = dpk_nlist_workspace_bytes(natoms, ncells);
  DPK_REQUIRE(need <= kNoallocScratchBytes, DPK_ERR_ARG,
              "dpk_nlist_build_noalloc: natoms=%d ncells=%d need %zu bytes of scratch, above the "
              "%zu-byte static scratch (use dpk_nlist_build with a workspace)",
              natoms, ncells, need, kNoallocScratchBytes);
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != DPK_OK) return rc;
  return dpk_nlist_build(nlist, coord_wrapped, status_dev, coord, atype, frame_of, frames,
                         nframes, natoms, ncells, img_pos, shift_vec, rcut, sel_host, ntypes,
                         max_cand, ws, kNoallocScratchBytes, stream);
}

extern "C" int dpk_nlist_pair_exclude(int* nlist, const int* atype, const int* frame_of,
                                      const dpk_frame* frames, int natoms, int nnei,
                                      const int* pair_mask, int ntypes, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && nnei > 0 && ntypes > 0 && ntypes <= kMaxTypes, DPK_ERR_ARG,
              "dpk_nlist_pair_exclude: bad sizes natoms=%d nnei=%d ntypes=%d", natoms, nnei,
              ntypes);
  DPK_REQUIRE(pair_mask != nullptr, DPK_ERR_ARG, "dpk_nlist_pair_exclude: pair_mask is NULL");
  if (natoms == 0) return DPK_OK;
  hipLaunchKernelGGL(k_pair_exclude, dim3(cdiv((long long)natoms * nnei, 256)), dim3(256), 0,
                     as_stream(stream), nlist, atype, frame_of, frames, natoms, nnei, pair_mask,
                     ntypes);
  DPK_LAUNCHED();
  return DPK_OK;
}
