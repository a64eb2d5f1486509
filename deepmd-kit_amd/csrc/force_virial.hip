// Forces and virials as the adjoint of the env matrix.
//
// Op forms (torch ops prod_force_se_a / prod_virial_se_a and their grads):
//   source/lib/src/prod_force.cc:23-73       F_i -= g.dR, F_j += g.dR
//   source/lib/src/prod_virial.cc:22-69      V[a][b] += (g.dR)[a] * rij[b]
//   source/lib/src/prod_force_grad.cc:22-76  d(force)/d(net_deriv) applied to grad
//   source/lib/src/prod_virial_grad.cc:22-63 d(virial)/d(net_deriv) applied to grad
// One wavefront per centre atom: the centre's own sum is reduced across the
// wavefront (one atomic per component instead of one per slot); neighbour
// contributions are FP64 atomics.  Virials are reduced per atom into a
// scratch row, then per frame in a fixed order (deterministic).
//
// Fused form (DeepPot path): the env-matrix derivative is recomputed from the
// FP64 coordinates (implicit images) instead of being read from a stored
// em_deriv (132 B/slot of traffic saved); neighbour forces go straight to the
// owning local atom (the ghost -> local reduction of
// deepmd/pt/model/model/transform_output.py:367-433).
#include "dpk_common.h"
#include "nlist_ws.h"

namespace dpk {
namespace {

__global__ void k_zero_f64(double* __restrict__ p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

int zero(double* p, long long n, hipStream_t st) {
  if (n <= 0) return DPK_OK;
  hipLaunchKernelGGL(k_zero_f64, dim3(cdiv(n, 256)), dim3(256), 0, st, p, n);
  DPK_LAUNCHED();
  return DPK_OK;
}

// ------------------------------------------------------------- op forms ---

__global__ void __launch_bounds__(256) k_prod_force(double* __restrict__ force,
                                                    const double* __restrict__ nd,
                                                    const double* __restrict__ dr,
                                                    const int* __restrict__ nlist, int nrows,
                                                    int nloc, int nall, int nnei) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nrows) return;
  const int ff = row / nloc, ii = row % nloc;
  double* fo = force + (size_t)ff * nall * 3;
  double ci[3] = {0.0, 0.0, 0.0};
  for (int jj = lane; jj < nnei; jj += kWave) {
    const size_t s = (size_t)row * nnei + jj;
    const double* g = nd + s * 4;
    const double* d = dr + s * 12;
    double f[3];
    for (int k = 0; k < 3; ++k) f[k] = g[0] * d[k] + g[1] * d[3 + k] + g[2] * d[6 + k] + g[3] * d[9 + k];
    for (int k = 0; k < 3; ++k) ci[k] -= f[k];
    const int j = nlist[s];
    if (j >= 0 && j < nall)  // entries outside the extended rows add nothing
      for (int k = 0; k < 3; ++k) atomicAdd(fo + 3 * j + k, f[k]);
  }
  for (int k = 0; k < 3; ++k) {
    double v = wave_sum(ci[k]);
    if (lane == 0) atomicAdd(fo + 3 * ii + k, v);
  }
}

__global__ void __launch_bounds__(256) k_prod_virial(double* __restrict__ vrow,
                                                     double* __restrict__ atom_virial,
                                                     const double* __restrict__ nd,
                                                     const double* __restrict__ dr,
                                                     const double* __restrict__ rij,
                                                     const int* __restrict__ nlist, int nrows,
                                                     int nloc, int nall, int nnei) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= nrows) return;
  const int ff = row / nloc;
  double* av = atom_virial ? atom_virial + (size_t)ff * nall * 9 : nullptr;
  double acc[9];
  for (int k = 0; k < 9; ++k) acc[k] = 0.0;
  for (int jj = lane; jj < nnei; jj += kWave) {
    const size_t s = (size_t)row * nnei + jj;
    const int j = nlist[s];
    if (j < 0 || j >= nall) continue;
    const double* g = nd + s * 4;
    const double* d = dr + s * 12;
    const double* r = rij + s * 3;
    double f[3];
    for (int k = 0; k < 3; ++k) f[k] = g[0] * d[k] + g[1] * d[3 + k] + g[2] * d[6 + k] + g[3] * d[9 + k];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) {
        const double t = f[a] * r[b];
        acc[3 * a + b] += t;
        if (av != nullptr) atomicAdd(av + (size_t)9 * j + 3 * a + b, t);
      }
  }
  for (int k = 0; k < 9; ++k) {
    double v = wave_sum(acc[k]);
    if (lane == 0) vrow[(size_t)row * 9 + k] = v;
  }
}

// Per-frame fixed-order sum of per-row virials: one block per frame.
__global__ void __launch_bounds__(256) k_virial_rows(double* __restrict__ virial,
                                                     const double* __restrict__ vrow, int nloc) {
  __shared__ double red[256];
  const int ff = blockIdx.x;
  for (int k = 0; k < 9; ++k) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nloc; i += 256) s += vrow[((size_t)ff * nloc + i) * 9 + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) virial[(size_t)ff * 9 + k] = red[0];
    __syncthreads();
  }
}

__global__ void k_force_grad(double* __restrict__ gn, const double* __restrict__ grad,
                             const double* __restrict__ dr, const int* __restrict__ nlist,
                             int nframes, int nloc, int nnei, int ngrad) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nframes * nloc * nnei) return;
  const long long row = idx / nnei;
  const int ff = (int)(row / nloc);
  const int ii = (int)(row % nloc);
  const double* gi = grad + ((size_t)ff * ngrad + ii) * 3;
  const double* d = dr + idx * 12;
  const int j = nlist[idx];
  double gj[3] = {0.0, 0.0, 0.0};
  if (j >= 0) {
    const int jr = j < ngrad ? j : j % nloc;
    const double* p = grad + ((size_t)ff * ngrad + jr) * 3;
    gj[0] = p[0];
    gj[1] = p[1];
    gj[2] = p[2];
  }
  for (int c = 0; c < 4; ++c) {
    double v = 0.0;
    for (int k = 0; k < 3; ++k) v += (gj[k] - gi[k]) * d[3 * c + k];
    gn[idx * 4 + c] = v;
  }
}

__global__ void k_virial_grad(double* __restrict__ gn, const double* __restrict__ grad,
                              const double* __restrict__ dr, const double* __restrict__ rij,
                              const int* __restrict__ nlist, int nframes, int nloc, int nnei) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nframes * nloc * nnei) return;
  const int ff = (int)(idx / ((long long)nloc * nnei));
  const double* G = grad + (size_t)ff * 9;
  const double* d = dr + idx * 12;
  const double* r = rij + idx * 3;
  const bool valid = nlist[idx] >= 0;
  for (int c = 0; c < 4; ++c) {
    double v = 0.0;
    if (valid)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) v += G[3 * a + b] * r[b] * d[3 * c + a];
    gn[idx * 4 + c] = v;
  }
}

// ----------------------------------------------------------- fused form ---

template <typename GT>
__global__ void __launch_bounds__(256) k_force_virial_fused(
    double* __restrict__ force, double* __restrict__ vrow, double* __restrict__ atom_virial,
    const GT* __restrict__ nd, const double* __restrict__ wrapped, const int* __restrict__ atype,
    const int* __restrict__ frame_of, const dpk_frame* __restrict__ frames,
    const int* __restrict__ nlist, const double* __restrict__ shift_vec,
    const double* __restrict__ stdv, int natoms, int nnei, double rcut, double rcut_smth,
    double protection, const int* __restrict__ pair_mask, int ntypes) {
  const int g = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= natoms) return;
  const int traw = atype[g];
  const int ti = traw < 0 ? 0 : traw;
  const int* pm = pair_mask ? pair_mask + (traw < 0 ? ntypes : traw) * (ntypes + 1) : nullptr;
  const dpk_frame& fd = frames[frame_of[g]];
  FastDiv byloc;
  byloc.init((unsigned)fd.nloc);
  const double cx = wrapped[3 * g], cy = wrapped[3 * g + 1], cz = wrapped[3 * g + 2];
  double ci[3] = {0.0, 0.0, 0.0};
  double vacc[9];
  for (int k = 0; k < 9; ++k) vacc[k] = 0.0;
  for (int jj = lane; jj < nnei; jj += kWave) {
    const size_t s = (size_t)g * nnei + jj;
    const int e = nlist[s];
    if (e < 0) continue;
    const int im = (int)byloc.div((unsigned)e);
    const int j = fd.atom_off + (e - im * fd.nloc);
    if (pm != nullptr && !pm[atype[j]]) continue;  // excluded pair: constant zero row
    const double* sv = shift_vec + 3 * (size_t)(fd.img_off + im);
    const double x = (wrapped[3 * j] + sv[0]) - cx;
    const double y = (wrapped[3 * j + 1] + sv[1]) - cy;
    const double z = (wrapped[3 * j + 2] + sv[2]) - cz;
    const double r = norm3_ref(x, y, z);
    double sw, dsw;
    spline5<double>(r, rcut_smth, rcut, sw, dsw);
    const double inr = 1.0 / r;
    const double iu = protection == 0.0 ? inr : 1.0 / (r + protection);
    const double dd[3] = {x, y, z};
    const double* sd = stdv + (size_t)ti * nnei * 4 + jj * 4;
    double gg[4];
    for (int c = 0; c < 4; ++c) gg[c] = (double)nd[s * 4 + c] / sd[c];
    double f[3];
    env_centre_deriv<double>(dd, inr, iu, sw, dsw, gg, f);
    for (int a = 0; a < 3; ++a) {
      ci[a] -= f[a];
      atomicAdd(force + 3 * (size_t)j + a, f[a]);
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) vacc[3 * a + b] += f[a] * dd[b];
    if (atom_virial != nullptr)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) atomicAdd(atom_virial + 9 * (size_t)j + 3 * a + b, f[a] * dd[b]);
  }
  for (int a = 0; a < 3; ++a) {
    double t = wave_sum(ci[a]);
    if (lane == 0) atomicAdd(force + 3 * (size_t)g + a, t);
  }
  for (int k = 0; k < 9; ++k) {
    double t = wave_sum(vacc[k]);
    if (lane == 0) vrow[(size_t)g * 9 + k] = t;
  }
}

// Cell-blocked form: one workgroup per cell of the neighbour-list grid, its
// wavefronts over the cell's atoms.  Every neighbour of these atoms lies in the
// 3x3x3 block of cells around it (cell edge >= rcut); its force contribution
// is accumulated in LDS at (relative cell, position inside that cell) -- exact
// integer keys from the binning and the image shift -- and the table is added
// to the global forces once per workgroup.  At C4 each neighbour receives ~7
// contributions per workgroup, so global FP64 atomics drop ~7x.  Contributions
// that fall outside the table (cells with more than kCellCap atoms, or
// rng > 1 in small boxes) use direct global atomics.
constexpr int kCellCap = 64;

__device__ __forceinline__ int floordiv_i(int a, int b) {
  const int q = a / b;
  return (a % b != 0 && ((a < 0) != (b < 0))) ? q - 1 : q;
}
constexpr int kCellTab = 27 * kCellCap;

template <typename GT>
__device__ __forceinline__ void slot_force(const GT* __restrict__ ndp, const double* __restrict__ sd,
                                           double x, double y, double z, double rcut,
                                           double rcut_smth, double protection, double f[3]) {
  // FP32 net_deriv: reciprocal-multiplies (within 1 ulp of the FP64
  // divisions, below the FP32 rounding of net_deriv); FP64 keeps divisions
  constexpr bool kExact = sizeof(GT) == 8;
  const double r = norm3_ref(x, y, z);
  double sw, dsw;
  spline5<double>(r, rcut_smth, rcut, sw, dsw);
  const double inr = kExact ? 1.0 / r : rcp_nr(r);
  const double iu = protection == 0.0 ? inr
                                      : (kExact ? 1.0 / (r + protection) : rcp_nr(r + protection));
  const double dd[3] = {x, y, z};
  double gg[4];
  for (int c = 0; c < 4; ++c) gg[c] = kExact ? (double)ndp[c] / sd[c] : (double)ndp[c] * rcp_nr(sd[c]);
  env_centre_deriv<double>(dd, inr, iu, sw, dsw, gg, f);
}

template <typename GT, int W>
__global__ void __launch_bounds__(64 * W) k_force_cells(
    double* __restrict__ force, double* __restrict__ vrow, double* __restrict__ atom_virial,
    const GT* __restrict__ nd, const double* __restrict__ wrapped, const int* __restrict__ atype,
    const int* __restrict__ frame_of, const dpk_frame* __restrict__ frames,
    const int* __restrict__ nlist, const double* __restrict__ shift_vec,
    const int* __restrict__ img_shift, const double* __restrict__ stdv,
    const int* __restrict__ cell_start, const int* __restrict__ cell_atoms,
    const int* __restrict__ atom_cell, const int* __restrict__ atom_slot, int nnei, double rcut,
    double rcut_smth, double protection, const int* __restrict__ pair_mask, int ntypes) {
  __shared__ double tab[kCellTab * 3];
  const int cell = blockIdx.x;
  const int a0 = cell_start[cell], a1 = cell_start[cell + 1];
  if (a1 <= a0) return;  // uniform over the workgroup
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int q = threadIdx.x; q < kCellTab * 3; q += blockDim.x) tab[q] = 0.0;
  const dpk_frame& fd = frames[frame_of[cell_atoms[a0]]];
  const int loc = cell - fd.cell_off;
  const int uc2 = loc % fd.nc[2], uc1 = (loc / fd.nc[2]) % fd.nc[1],
            uc0 = loc / (fd.nc[2] * fd.nc[1]);
  FastDiv byloc, bync2, bync1;  // per-slot decoding of e and of the cell id
  byloc.init((unsigned)fd.nloc);
  bync2.init((unsigned)fd.nc[2]);
  bync1.init((unsigned)fd.nc[1]);
  __syncthreads();
  for (int ai = a0 + wave; ai < a1; ai += W) {
    const int g = cell_atoms[ai];
    const int traw = atype[g];
    const int ti = traw < 0 ? 0 : traw;
    const int* pm = pair_mask ? pair_mask + (traw < 0 ? ntypes : traw) * (ntypes + 1) : nullptr;
    const double cx = wrapped[3 * g], cy = wrapped[3 * g + 1], cz = wrapped[3 * g + 2];
    double ci[3] = {0.0, 0.0, 0.0};
    double vacc[9];
    for (int k = 0; k < 9; ++k) vacc[k] = 0.0;
    for (int jj = lane; jj < nnei; jj += kWave) {
      const size_t s = (size_t)g * nnei + jj;
      const int e = nlist[s];
      if (e < 0) continue;
      const int im = (int)byloc.div((unsigned)e);
      const int j = fd.atom_off + (e - im * fd.nloc);
      if (pm != nullptr && !pm[atype[j]]) continue;  // excluded pair: constant zero row
      const int gi = fd.img_off + im;
      const double* sv = shift_vec + 3 * (size_t)gi;
      const double x = (wrapped[3 * j] + sv[0]) - cx;
      const double y = (wrapped[3 * j + 1] + sv[1]) - cy;
      const double z = (wrapped[3 * j + 2] + sv[2]) - cz;
      double f[3];
      slot_force<GT>(nd + s * 4, stdv + (size_t)ti * nnei * 4 + jj * 4, x, y, z, rcut, rcut_smth,
                     protection,
                     f);
      for (int a = 0; a < 3; ++a) ci[a] -= f[a];
      // neighbour's cell relative to this cell, image included
      const int lj = atom_cell[j] - fd.cell_off;
      const int q2 = (int)bync2.div((unsigned)lj), uj2 = lj - q2 * fd.nc[2];
      const int uj0 = (int)bync1.div((unsigned)q2), uj1 = q2 - uj0 * fd.nc[1];
      const int d0 = uj0 + img_shift[3 * gi] * fd.nc[0] - uc0;
      const int d1 = uj1 + img_shift[3 * gi + 1] * fd.nc[1] - uc1;
      const int d2 = uj2 + img_shift[3 * gi + 2] * fd.nc[2] - uc2;
      const int sl = atom_slot[j];
      if (d0 >= -1 && d0 <= 1 && d1 >= -1 && d1 <= 1 && d2 >= -1 && d2 <= 1 && sl < kCellCap) {
        double* t = tab + 3 * ((((d0 + 1) * 3 + (d1 + 1)) * 3 + (d2 + 1)) * kCellCap + sl);
        for (int a = 0; a < 3; ++a) atomicAdd(t + a, f[a]);
      } else {
        for (int a = 0; a < 3; ++a) atomicAdd(force + 3 * (size_t)j + a, f[a]);
      }
      const double dd[3] = {x, y, z};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) vacc[3 * a + b] += f[a] * dd[b];
      if (atom_virial != nullptr) {
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            atomicAdd(atom_virial + 9 * (size_t)j + 3 * a + b, f[a] * dd[b]);
      }
    }
    for (int a = 0; a < 3; ++a) {
      const double t = wave_sum(ci[a]);
      if (lane == 0) {
        const int sl = atom_slot[g];
        if (sl < kCellCap) atomicAdd(tab + 3 * (13 * kCellCap + sl) + a, t);
        else atomicAdd(force + 3 * (size_t)g + a, t);
      }
    }
    for (int k = 0; k < 9; ++k) {
      const double t = wave_sum(vacc[k]);
      if (lane == 0) vrow[(size_t)g * 9 + k] = t;
    }
  }
  __syncthreads();
  // flush: entry (relative cell, slot) -> that cell's atom
  for (int q = threadIdx.x; q < kCellTab; q += blockDim.x) {
    const double fx = tab[3 * q], fy = tab[3 * q + 1], fz = tab[3 * q + 2];
    if (fx == 0.0 && fy == 0.0 && fz == 0.0) continue;
    const int rc = q / kCellCap, sl = q - rc * kCellCap;
    const int d0 = rc / 9 - 1, d1 = (rc / 3) % 3 - 1, d2 = rc % 3 - 1;
    int u0 = uc0 + d0, u1 = uc1 + d1, u2 = uc2 + d2;
    // wrap back into the grid (the image only changed the position)
    u0 -= floordiv_i(u0, fd.nc[0]) * fd.nc[0];
    u1 -= floordiv_i(u1, fd.nc[1]) * fd.nc[1];
    u2 -= floordiv_i(u2, fd.nc[2]) * fd.nc[2];
    const int cid = fd.cell_off + (u0 * fd.nc[1] + u1) * fd.nc[2] + u2;
    const int j = cell_atoms[cell_start[cid] + sl];
    atomicAdd(force + 3 * (size_t)j, fx);
    atomicAdd(force + 3 * (size_t)j + 1, fy);
    atomicAdd(force + 3 * (size_t)j + 2, fz);
  }
}

// Per-frame sums of per-atom rows (frames are contiguous atom ranges), in a
// fixed order: chunk partials (one workgroup per kFrameChunk atoms of a
// frame), then one workgroup per frame over its partials.  Deterministic for
// a given (frames, max_nloc), unlike atomics into one address per frame,
// which also serialise (131k FP64 atomics on one address take ~10 ms).
constexpr int kFrameChunk = 2048;

template <int NC>
__device__ __forceinline__ void block_sum_store(double (&v)[NC], double* out) {
  __shared__ double red[4][NC];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < NC; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NC; ++k) red[wave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < NC) {
    const int k = threadIdx.x;
    out[k] = ((red[0][k] + red[1][k]) + (red[2][k] + red[3][k]));
  }
}

template <int NC>
__global__ void __launch_bounds__(256) k_frame_partial(double* __restrict__ part,
                                                       const double* __restrict__ in,
                                                       const dpk_frame* __restrict__ frames,
                                                       int nchunk) {
  const int f = blockIdx.y, x = blockIdx.x;
  const dpk_frame& fd = frames[f];
  const int i0 = x * kFrameChunk;
  const int i1 = min(fd.nloc, i0 + kFrameChunk);
  double v[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) v[k] = 0.0;
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
    const double* r = in + ((size_t)fd.atom_off + i) * NC;
#pragma unroll
    for (int k = 0; k < NC; ++k) v[k] += r[k];
  }
  block_sum_store<NC>(v, part + ((size_t)f * nchunk + x) * NC);
}

template <int NC>
__global__ void __launch_bounds__(256) k_frame_final(double* __restrict__ out,
                                                     const double* __restrict__ part,
                                                     int nchunk) {
  const int f = blockIdx.x;
  double v[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) v[k] = 0.0;
  for (int x = threadIdx.x; x < nchunk; x += 256) {
    const double* r = part + ((size_t)f * nchunk + x) * NC;
#pragma unroll
    for (int k = 0; k < NC; ++k) v[k] += r[k];
  }
  block_sum_store<NC>(v, out + (size_t)f * NC);
}

template <int NC>
int frame_sum_t(double* out, const double* in, const dpk_frame* frames, int nframes, int nchunk,
                double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_frame_partial<NC>, dim3(nchunk, nframes), dim3(256), 0, st, part, in,
                     frames, nchunk);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_frame_final<NC>, dim3(nframes), dim3(256), 0, st, out, part, nchunk);
  DPK_LAUNCHED();
  return DPK_OK;
}


// ------------------------------------------------- gather form (MD path) --
// For an engine's extended region (explicit ghosts: the list holds
// extended-atom indices), neighbour forces are gathered instead of added
// with atomics: after each formatting the slots are binned by neighbour into
// fixed-capacity buckets (cap ~ the slot count: an atom is a neighbour of
// about as many centres as it has neighbours; slots beyond a full bucket go
// to an overflow list, added with atomics at the end), one pass writes
// every slot's force share f (summing the centre's -f and virial row in
// registers), a second pass sums each extended atom's bucket.

__global__ void k_bucket_zero(int* __restrict__ count, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q <= n) count[q] = 0;  // count[n]: the overflow length
}

// the slot's neighbour (or -1: empty, outside the region, excluded pair)
__device__ __forceinline__ int gather_nbr(const int* __restrict__ nlist,
                                          const int* __restrict__ atype,
                                          const int* __restrict__ pair_mask, int ntypes,
                                          long long s, int nnei, int nall) {
  const int e = nlist[s];
  if (e < 0 || e >= nall) return -1;
  if (pair_mask != nullptr) {
    const int traw = atype[(int)(s / nnei)];
    const int* pm = pair_mask + (traw < 0 ? ntypes : traw) * (ntypes + 1);
    const int tj = atype[e];
    if (!pm[tj < 0 ? ntypes : tj]) return -1;
  }
  return e;
}

__global__ void k_bucket_fill(int* __restrict__ bucket, int* __restrict__ count,
                              int* __restrict__ ovf, const int* __restrict__ nlist,
                              const int* __restrict__ atype, const int* __restrict__ pair_mask,
                              int ntypes, long long total, int nnei, int nall, int cap) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= total) return;
  const int e = gather_nbr(nlist, atype, pair_mask, ntypes, s, nnei, nall);
  if (e < 0) return;
  const int pos = atomicAdd(count + e, 1);
  if (pos < cap) bucket[(size_t)e * cap + pos] = (int)s;
  else ovf[atomicAdd(count + nall, 1)] = (int)s;
}

// One wavefront per centre: fs[slot] = the force share of the slot's
// neighbour, force[centre] = -sum of them, vrow[centre] = sum f (x) d.
template <typename GT>
__global__ void __launch_bounds__(256) k_slot_forces(
    double* __restrict__ fs, double* __restrict__ force, double* __restrict__ vrow,
    const GT* __restrict__ nd, const double* __restrict__ coord, const int* __restrict__ atype,
    const int* __restrict__ nlist, const double* __restrict__ stdv, int natoms, int nnei,
    int nall, double rcut, double rcut_smth, double protection,
    const int* __restrict__ pair_mask, int ntypes) {
  const int g = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= natoms) return;
  const int traw = atype[g];
  const int ti = traw < 0 ? 0 : traw;
  const double cx = coord[3 * g], cy = coord[3 * g + 1], cz = coord[3 * g + 2];
  double ci[3] = {0.0, 0.0, 0.0};
  double vacc[9];
  for (int k = 0; k < 9; ++k) vacc[k] = 0.0;
  for (int jj = lane; jj < nnei; jj += kWave) {
    const size_t s = (size_t)g * nnei + jj;
    const int j = gather_nbr(nlist, atype, pair_mask, ntypes, (long long)s, nnei, nall);
    double f[3] = {0.0, 0.0, 0.0};
    if (j >= 0) {
      const double dd[3] = {coord[3 * j] - cx, coord[3 * j + 1] - cy, coord[3 * j + 2] - cz};
      const double r = norm3_ref(dd[0], dd[1], dd[2]);
      double sw, dsw;
      spline5<double>(r, rcut_smth, rcut, sw, dsw);
      const double inr = 1.0 / r;
      const double iu = protection == 0.0 ? inr : 1.0 / (r + protection);
      const double* sd = stdv + (size_t)ti * nnei * 4 + jj * 4;
      double gg[4];
      for (int c = 0; c < 4; ++c) gg[c] = (double)nd[s * 4 + c] / sd[c];
      env_centre_deriv<double>(dd, inr, iu, sw, dsw, gg, f);
      for (int a = 0; a < 3; ++a) {
        ci[a] -= f[a];
        for (int b = 0; b < 3; ++b) vacc[3 * a + b] += f[a] * dd[b];
      }
    }
    for (int a = 0; a < 3; ++a) fs[3 * s + a] = f[a];
  }
  for (int a = 0; a < 3; ++a) {
    double t = wave_sum(ci[a]);
    if (lane == 0) force[3 * (size_t)g + a] = t;
  }
  for (int k = 0; k < 9; ++k) {
    double t = wave_sum(vacc[k]);
    if (lane == 0) vrow[(size_t)g * 9 + k] = t;
  }
}

// One wavefront per extended atom j: force[j] (= its centre part when j is
// local, else 0) + the sum of its bucket's shares; atom_virial[j]
// (optional) = sum f (x) (x_j - x_centre) over them.
__global__ void __launch_bounds__(256) k_gather_forces(
    double* __restrict__ force, double* __restrict__ atom_virial, const double* __restrict__ fs,
    const int* __restrict__ count, const int* __restrict__ bucket,
    const double* __restrict__ coord, int natoms, int nnei, int nall, int cap) {
  const int j = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= nall) return;
  const int n = count[j] < cap ? count[j] : cap;
  const int* row = bucket + (size_t)j * cap;
  double f[3] = {0.0, 0.0, 0.0}, v[9];
  for (int k = 0; k < 9; ++k) v[k] = 0.0;
  const double xj = coord[3 * j], yj = coord[3 * j + 1], zj = coord[3 * j + 2];
  for (int q = lane; q < n; q += kWave) {
    const int s = row[q];
    const double fx = fs[3 * (size_t)s], fy = fs[3 * (size_t)s + 1], fz = fs[3 * (size_t)s + 2];
    f[0] += fx;
    f[1] += fy;
    f[2] += fz;
    if (atom_virial != nullptr) {
      const int i = s / nnei;
      const double d[3] = {xj - coord[3 * i], yj - coord[3 * i + 1], zj - coord[3 * i + 2]};
      const double ff[3] = {fx, fy, fz};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) v[3 * a + b] += ff[a] * d[b];
    }
  }
  for (int a = 0; a < 3; ++a) {
    const double t = wave_sum(f[a]);
    if (lane == 0) force[3 * (size_t)j + a] = (j < natoms ? force[3 * (size_t)j + a] : 0.0) + t;
  }
  if (atom_virial != nullptr)
    for (int k = 0; k < 9; ++k) {
      const double t = wave_sum(v[k]);
      if (lane == 0) atom_virial[9 * (size_t)j + k] = t;
    }
}

// Slots that did not fit their bucket (count[nall] of them): added with
// atomics, grid-stride over the device-side count.
__global__ void k_gather_overflow(double* __restrict__ force, double* __restrict__ atom_virial,
                                  const double* __restrict__ fs, const int* __restrict__ count,
                                  const int* __restrict__ ovf, const int* __restrict__ nlist,
                                  const double* __restrict__ coord, int nnei, int nall) {
  const int n = count[nall];
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int s = ovf[q];
    const int j = nlist[s], i = s / nnei;
    const double ff[3] = {fs[3 * (size_t)s], fs[3 * (size_t)s + 1], fs[3 * (size_t)s + 2]};
    for (int a = 0; a < 3; ++a) atomicAdd(force + 3 * (size_t)j + a, ff[a]);
    if (atom_virial != nullptr) {
      const double d[3] = {coord[3 * j] - coord[3 * i], coord[3 * j + 1] - coord[3 * i + 1],
                           coord[3 * j + 2] - coord[3 * i + 2]};
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) atomicAdd(atom_virial + 9 * (size_t)j + 3 * a + b, ff[a] * d[b]);
    }
  }
}

}  // namespace
}  // namespace dpk

using namespace dpk;

extern "C" int dpk_prod_force_a_f64(double* force, const double* net_deriv,
                                    const double* in_deriv, const int* nlist, int nframes,
                                    int nloc, int nall, int nnei, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nframes >= 0 && nloc >= 0 && nall >= nloc && nnei > 0, DPK_ERR_ARG,
              "prod_force_a: bad sizes");
  hipStream_t st = as_stream(stream);
  int rc = zero(force, (long long)nframes * nall * 3, st);
  if (rc) return rc;
  int nrows = nframes * nloc;
  if (nrows == 0) return DPK_OK;
  hipLaunchKernelGGL(k_prod_force, dim3(cdiv(nrows, 4)), dim3(256), 0, st, force, net_deriv,
                     in_deriv, nlist, nrows, nloc, nall, nnei);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" size_t dpk_prod_virial_workspace_bytes(int nframes, int nloc) {
  return sizeof(double) * 9 * (size_t)nframes * (size_t)nloc;
}

extern "C" int dpk_prod_virial_a_f64(double* virial, double* atom_virial,
                                     const double* net_deriv, const double* in_deriv,
                                     const double* rij, const int* nlist, int nframes, int nloc,
                                     int nall, int nnei, void* workspace, size_t workspace_bytes,
                                     dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nframes >= 0 && nloc >= 0 && nall >= nloc && nnei > 0, DPK_ERR_ARG,
              "prod_virial_a: bad sizes");
  DPK_REQUIRE(workspace_bytes >= dpk_prod_virial_workspace_bytes(nframes, nloc), DPK_ERR_ARG,
              "prod_virial_a: workspace too small");
  hipStream_t st = as_stream(stream);
  int rc = atom_virial ? zero(atom_virial, (long long)nframes * nall * 9, st) : DPK_OK;
  if (rc) return rc;
  rc = zero(virial, (long long)nframes * 9, st);
  if (rc) return rc;
  int nrows = nframes * nloc;
  if (nrows == 0) return DPK_OK;
  double* vrow = static_cast<double*>(workspace);
  hipLaunchKernelGGL(k_prod_virial, dim3(cdiv(nrows, 4)), dim3(256), 0, st, vrow, atom_virial,
                     net_deriv, in_deriv, rij, nlist, nrows, nloc, nall, nnei);
  DPK_LAUNCHED();
  hipLaunchKernelGGL(k_virial_rows, dim3(nframes), dim3(256), 0, st, virial, vrow, nloc);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_prod_force_grad_a_f64(double* grad_net, const double* grad,
                                         const double* in_deriv, const int* nlist, int nframes,
                                         int nloc, int nnei, int ngrad, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nframes >= 0 && nloc >= 0 && nnei > 0 && ngrad >= nloc, DPK_ERR_ARG,
              "prod_force_grad_a: bad sizes");
  long long total = (long long)nframes * nloc * nnei;
  if (total == 0) return DPK_OK;
  hipLaunchKernelGGL(k_force_grad, dim3(cdiv(total, 256)), dim3(256), 0, as_stream(stream),
                     grad_net, grad, in_deriv, nlist, nframes, nloc, nnei, ngrad);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_prod_virial_grad_a_f64(double* grad_net, const double* grad,
                                          const double* in_deriv, const double* rij,
                                          const int* nlist, int nframes, int nloc, int nnei,
                                          dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nframes >= 0 && nloc >= 0 && nnei > 0, DPK_ERR_ARG, "prod_virial_grad_a: bad sizes");
  long long total = (long long)nframes * nloc * nnei;
  if (total == 0) return DPK_OK;
  hipLaunchKernelGGL(k_virial_grad, dim3(cdiv(total, 256)), dim3(256), 0, as_stream(stream),
                     grad_net, grad, in_deriv, rij, nlist, nframes, nloc, nnei);
  DPK_LAUNCHED();
  return DPK_OK;
}

// Workspace of dpk_force_virial_gather: count (nall + 1 ints, the last the
// overflow length), buckets (nall * cap ints), the overflow list and the
// force shares (natoms * nnei ints / doubles x 3); cap = nnei rounded up to 64.
static int gather_cap(int nnei) { return (nnei + 63) / 64 * 64; }
static size_t gather_ws_bytes(int natoms, int nnei, int nall, size_t* off_b, size_t* off_o,
                              size_t* off_fs) {
  const size_t slots = (size_t)natoms * (size_t)nnei;
  auto al = [](size_t o) { return (o + 255) & ~(size_t)255; };
  size_t o = al(sizeof(int) * ((size_t)nall + 1));
  if (off_b) *off_b = o;
  o = al(o + sizeof(int) * (size_t)nall * gather_cap(nnei));
  if (off_o) *off_o = o;
  o = al(o + sizeof(int) * slots);
  if (off_fs) *off_fs = o;
  return o + sizeof(double) * 3 * slots;
}

extern "C" size_t dpk_force_virial_gather_workspace_bytes(int natoms, int nnei, int nall) {
  if (natoms < 0 || nnei <= 0 || nall < natoms) return 0;
  return gather_ws_bytes(natoms, nnei, nall, nullptr, nullptr, nullptr);
}

extern "C" int dpk_force_virial_gather(double* force, double* atom_virial, double* vrow,
                                       const void* net_deriv, int prec, const double* coord,
                                       const int* atype, const int* nlist, const double* std,
                                       int natoms, int nnei, int nall, double rcut,
                                       double rcut_smth, double protection, const int* pair_mask,
                                       int ntypes, void* workspace, size_t workspace_bytes,
                                       dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && nnei > 0 && nall >= natoms, DPK_ERR_ARG,
              "force_virial_gather: bad sizes natoms=%d nnei=%d nall=%d", natoms, nnei, nall);
  DPK_REQUIRE(protection >= 0.0, DPK_ERR_ARG, "force_virial_gather: protection < 0");
  DPK_REQUIRE(prec == DPK_PREC_F64 || prec == DPK_PREC_F32 || prec == DPK_PREC_BF16 ||
                  prec == DPK_PREC_F16 || prec == DPK_PREC_F16X2,
              DPK_ERR_ARG, "force_virial_gather: unknown precision %d", prec);
  size_t off_b = 0, off_o = 0, off_fs = 0;
  const size_t need = gather_ws_bytes(natoms, nnei, nall, &off_b, &off_o, &off_fs);
  DPK_REQUIRE(workspace != nullptr && workspace_bytes >= need, DPK_ERR_ARG,
              "force_virial_gather: workspace %zu < %zu bytes", workspace_bytes, need);
  DPK_REQUIRE((long long)natoms * nnei < (1ll << 31), DPK_ERR_UNSUPPORTED,
              "force_virial_gather: more than 2^31 slots");
  if (nall == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  unsigned char* ws = static_cast<unsigned char*>(workspace);
  int* count = reinterpret_cast<int*>(ws);
  int* bucket = reinterpret_cast<int*>(ws + off_b);
  int* ovf = reinterpret_cast<int*>(ws + off_o);
  double* fs = reinterpret_cast<double*>(ws + off_fs);
  const int cap = gather_cap(nnei);
  const long long total = (long long)natoms * nnei;
  hipLaunchKernelGGL(k_bucket_zero, dim3(cdiv((long long)nall + 1, 256)), dim3(256), 0, st, count,
                     nall);
  if (total)
    hipLaunchKernelGGL(k_bucket_fill, dim3(cdiv(total, 256)), dim3(256), 0, st, bucket, count,
                       ovf, nlist, atype, pair_mask, ntypes, total, nnei, nall, cap);
  if (natoms) {
    if (prec == DPK_PREC_F64)
      hipLaunchKernelGGL(k_slot_forces<double>, dim3(cdiv(natoms, 4)), dim3(256), 0, st, fs,
                         force, vrow, static_cast<const double*>(net_deriv), coord, atype, nlist,
                         std, natoms, nnei, nall, rcut, rcut_smth, protection, pair_mask, ntypes);
    else
      hipLaunchKernelGGL(k_slot_forces<float>, dim3(cdiv(natoms, 4)), dim3(256), 0, st, fs,
                         force, vrow, static_cast<const float*>(net_deriv), coord, atype, nlist,
                         std, natoms, nnei, nall, rcut, rcut_smth, protection, pair_mask, ntypes);
  }
  hipLaunchKernelGGL(k_gather_forces, dim3(cdiv(nall, 4)), dim3(256), 0, st, force, atom_virial,
                     fs, count, bucket, coord, natoms, nnei, nall, cap);
  hipLaunchKernelGGL(k_gather_overflow, dim3(64), dim3(256), 0, st, force, atom_virial, fs, count,
                     ovf, nlist, coord, nnei, nall);
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" size_t dpk_force_virial_workspace_bytes(int natoms) {
  return sizeof(double) * 9 * (size_t)natoms;
}

extern "C" int dpk_force_virial_fused(double* force, double* atom_virial, const void* net_deriv,
                                      int prec, const double* coord_wrapped, const int* atype,
                                      const int* frame_of, const dpk_frame* frames,
                                      const int* nlist, const double* shift_vec,
                                      const double* std, int natoms, int nnei, double rcut,
                                      double rcut_smth, double protection, void* workspace,
                                      size_t workspace_bytes, const int* pair_mask, int ntypes,
                                      dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && nnei > 0, DPK_ERR_ARG, "force_virial_fused: bad sizes");
  DPK_REQUIRE(protection >= 0.0, DPK_ERR_ARG, "force_virial_fused: protection < 0");
  DPK_REQUIRE(workspace_bytes >= dpk_force_virial_workspace_bytes(natoms), DPK_ERR_ARG,
              "force_virial_fused: workspace too small");
  if (natoms == 0) return DPK_OK;
  hipStream_t st = as_stream(stream);
  double* vrow = static_cast<double*>(workspace);
  dim3 grid(cdiv(natoms, 4)), block(256);
  if (prec == DPK_PREC_F64) {
    hipLaunchKernelGGL(k_force_virial_fused<double>, grid, block, 0, st, force, vrow, atom_virial,
                       static_cast<const double*>(net_deriv), coord_wrapped, atype, frame_of,
                       frames, nlist, shift_vec, std, natoms, nnei, rcut, rcut_smth, protection,
                       pair_mask, ntypes);
  } else if (prec == DPK_PREC_F32 || prec == DPK_PREC_BF16 || prec == DPK_PREC_F16 ||
             prec == DPK_PREC_F16X2) {
    hipLaunchKernelGGL(k_force_virial_fused<float>, grid, block, 0, st, force, vrow, atom_virial,
                       static_cast<const float*>(net_deriv), coord_wrapped, atype, frame_of,
                       frames, nlist, shift_vec, std, natoms, nnei, rcut, rcut_smth, protection,
                       pair_mask, ntypes);
  } else {
    set_error("force_virial_fused: unknown precision %d", prec);
    return DPK_ERR_ARG;
  }
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" int dpk_force_virial_cells(double* force, double* atom_virial, const void* net_deriv,
                                      int prec, const double* coord_wrapped, const int* atype,
                                      const int* frame_of, const dpk_frame* frames,
                                      const int* nlist, const double* shift_vec,
                                      const int* img_shift, const double* std, int natoms,
                                      int ncells, int nnei, double rcut, double rcut_smth,
                                      double protection, void* nlist_workspace,
                                      size_t nlist_workspace_bytes,
                                      void* workspace, size_t workspace_bytes,
                                      const int* pair_mask, int ntypes, dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(natoms >= 0 && nnei > 0 && ncells > 0, DPK_ERR_ARG, "force_virial_cells: bad sizes");
  DPK_REQUIRE(protection >= 0.0, DPK_ERR_ARG, "force_virial_cells: protection < 0");
  DPK_REQUIRE(workspace_bytes >= dpk_force_virial_workspace_bytes(natoms), DPK_ERR_ARG,
              "force_virial_cells: workspace too small");
  DPK_REQUIRE(nlist_workspace && nlist_workspace_bytes >= nlist_ws_bytes(natoms, ncells),
              DPK_ERR_ARG, "force_virial_cells: the neighbour-list workspace is required");
  if (natoms == 0) return DPK_OK;
  if (tune(DPK_TUNE_FORCE_MODE) == 1)
    return dpk_force_virial_fused(force, atom_virial, net_deriv, prec, coord_wrapped, atype,
                                  frame_of, frames, nlist, shift_vec, std, natoms, nnei, rcut,
                                  rcut_smth, protection, workspace, workspace_bytes, pair_mask,
                                  ntypes, stream);
  hipStream_t st = as_stream(stream);
  double* vrow = static_cast<double*>(workspace);
  const NlistWs w = nlist_ws(nlist_workspace, natoms, ncells);
  DPK_REQUIRE(prec == DPK_PREC_F64 || prec == DPK_PREC_F32 || prec == DPK_PREC_BF16 ||
                  prec == DPK_PREC_F16 || prec == DPK_PREC_F16X2,
              DPK_ERR_ARG, "force_virial_cells: unknown precision %d", prec);
  // wavefronts per cell workgroup (sharing one LDS table): from the atoms per
  // cell unless set (DPK_TUNE_FORCE_WAVES).  16 only for very full cells: at
  // ~48 atoms per cell (FCC Cu, rcut 6) 8 wavefronts take 530 us where 16
  // take 625 us (two 41 KB workgroups of 16 fill a CU's 32 wave slots; three
  // of 8 leave room for the next cell's start-up)
  int nw = tune(DPK_TUNE_FORCE_WAVES);
  if (nw == 0) nw = natoms >= 96 * ncells ? 16 : (natoms >= 16 * ncells ? 8 : 4);
  nw = nw <= 4 ? 4 : (nw <= 8 ? 8 : 16);
#define DPK_FORCE_CELLS(GT, NW)                                                              \
  hipLaunchKernelGGL((k_force_cells<GT, NW>), dim3(ncells), dim3(64 * NW), 0, st, force, vrow, \
                     atom_virial, static_cast<const GT*>(net_deriv), coord_wrapped, atype,     \
                     frame_of, frames, nlist, shift_vec, img_shift, std, w.cell_start,         \
                     w.cell_atoms, w.atom_cell, w.atom_slot, nnei, rcut, rcut_smth,             \
                     protection, pair_mask, ntypes)
#define DPK_FORCE_CELLS_W(GT)            \
  if (nw == 16) DPK_FORCE_CELLS(GT, 16); \
  else if (nw == 8) DPK_FORCE_CELLS(GT, 8); \
  else DPK_FORCE_CELLS(GT, 4)
  if (prec == DPK_PREC_F64) {
    DPK_FORCE_CELLS_W(double);
  } else {
    DPK_FORCE_CELLS_W(float);
  }
#undef DPK_FORCE_CELLS_W
#undef DPK_FORCE_CELLS
  DPK_LAUNCHED();
  return DPK_OK;
}

extern "C" size_t dpk_frame_sum_workspace_bytes(int nframes, int max_nloc, int ncomp) {
  if (nframes <= 0 || max_nloc < 0 || ncomp <= 0) return 0;
  const size_t nchunk = (size_t)cdiv(max_nloc > 0 ? max_nloc : 1, kFrameChunk);
  return sizeof(double) * (size_t)nframes * nchunk * (size_t)ncomp;
}

extern "C" int dpk_frame_sum(double* out, const double* in, int ncomp, const dpk_frame* frames,
                             int nframes, int max_nloc, void* workspace, size_t workspace_bytes,
                             dpk_stream_t stream) {
  clear_error();
  DPK_REQUIRE(nframes > 0 && max_nloc >= 0, DPK_ERR_ARG, "frame_sum: nframes=%d max_nloc=%d",
              nframes, max_nloc);
  DPK_REQUIRE(ncomp == 1 || ncomp == 9, DPK_ERR_UNSUPPORTED, "frame_sum: ncomp=%d (1 or 9)",
              ncomp);
  DPK_REQUIRE(workspace_bytes >= dpk_frame_sum_workspace_bytes(nframes, max_nloc, ncomp),
              DPK_ERR_ARG, "frame_sum: workspace too small");
  const int nchunk = cdiv(max_nloc > 0 ? max_nloc : 1, kFrameChunk);
  double* part = static_cast<double*>(workspace);
  hipStream_t st = as_stream(stream);
  return ncomp == 1 ? frame_sum_t<1>(out, in, frames, nframes, nchunk, part, st)
                    : frame_sum_t<9>(out, in, frames, nframes, nchunk, part, st);
}
