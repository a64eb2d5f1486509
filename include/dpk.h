/*
 * dpk.h -- C ABI of libdpk, the MI355X (gfx950) kernels of the Deep Potential
 * se_e2_a energy / force / virial path.
 *
 * Conventions for every entry point:
 *   - returns DPK_OK (0) or a DPK_ERR_* code; dpk_last_error() gives the
 *     message of the last failure on the calling host thread;
 *   - array arguments are DEVICE pointers unless the name ends in "_host";
 *   - work is enqueued on `stream` (a hipStream_t); no entry point
 *     synchronises the device or allocates memory, so every call can be
 *     captured into a hipGraph; scratch memory is passed in explicitly;
 *   - no exceptions cross this boundary; callable from several host threads
 *     on distinct streams.
 *
 * Layouts follow the reference's flat layouts: nlist [nf, nloc, nnei] with
 * -1 padding, type blocks |sel0|sel1|..., extended index e = s * nloc + j
 * (image s in the order of the reference's shift list, local atom j);
 * em [nf, nloc, nnei*4]; em_deriv [nf, nloc, nnei*12] (d em / d r_centre);
 * rij [nf, nloc, nnei*3]; force [nf, nall*3]; virial [nf, 9].
 */
#ifndef DPK_H
#define DPK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* dpk_stream_t; /* hipStream_t */

enum {
  DPK_OK = 0,
  DPK_ERR_ARG = 1,         /* bad sizes / null pointers / unsupported option */
  DPK_ERR_HIP = 2,         /* a HIP runtime call failed */
  DPK_ERR_CAPACITY = 3,    /* a bounded buffer (neighbour candidates, ...) overflowed */
  DPK_ERR_UNSUPPORTED = 4  /* shape outside what the kernels are built for */
};

/* Message of the last error on this host thread ("" when none). */
const char* dpk_last_error(void);
/* ABI version of this header (bumped on incompatible change). */
int dpk_abi_version(void);

/*
 * Launch tunables.  No counterpart in the reference: they select among
 * equivalent kernel organisations whose speed depends on the system (chosen
 * by deepmd_amd.autotune after timing them).  A launch reads the calling
 * thread's current tuning set (dpk_tuning_use), else the process-wide
 * defaults (dpk_tune_set / dpk_tune_get).  A model owns one set and makes it
 * current around its launches, so models with different shapes, on any
 * streams or threads, keep their own choices; the results do not depend on
 * the tunables.
 */
enum {
  DPK_TUNE_NLIST_MODE = 0,      /* 0: staged cell neighbourhoods (default), 1: per-centre gather */
  DPK_TUNE_NLIST_STAGE_PCT = 1, /* staging capacity in % of the mean neighbourhood, 10..400 */
  DPK_TUNE_FORCE_MODE = 2,      /* 0: cell-blocked forces (default), 1: per-atom */
  DPK_TUNE_EMBED_GRID_PCT = 3,  /* BF16 embedding persistent grid, % of resident slots;
                                   0 = one workgroup per 4 atoms */
  DPK_TUNE_NLIST_WAVES = 4,     /* wavefronts per cell workgroup of the search: 0 = from the
                                   atoms per cell, else 4, 8 or 16 (rounded up) */
  DPK_TUNE_FORCE_WAVES = 5,     /* the same for the cell-blocked force kernel */
  DPK_TUNE_NLIST_SORT = 6,      /* 0: packed 64-bit (distance, index) keys with the full-key
                                   sort for rows holding (near-)equal distances (default),
                                   1: full-key sort for every row */
  DPK_TUNE_COUNT = 7
};
int dpk_tune_set(int key, int value); /* process-wide default */
int dpk_tune_get(int key);

typedef struct dpk_tuning {
  int v[DPK_TUNE_COUNT];
} dpk_tuning;
/* a set holding the current defaults (NULL on host allocation failure) */
dpk_tuning* dpk_tuning_create(void);
void dpk_tuning_destroy(dpk_tuning* t);
int dpk_tuning_set(dpk_tuning* t, int key, int value); /* range-checked as dpk_tune_set */
int dpk_tuning_get(const dpk_tuning* t, int key);      /* t NULL: the default */
/* make t the calling thread's current set (NULL: the defaults); returns the
 * previous one, for the caller to restore.  t must outlive its use. */
const dpk_tuning* dpk_tuning_use(const dpk_tuning* t);

/* ------------------------------------------------------------------------ *
 * Neighbour list: cell-list search + per-type format (PT semantics).
 * Replaces deepmd/pt/utils/nlist.py:14-199,304-337,424-513 (normalize_coord,
 * extend_coord_with_ghosts, build_neighbor_list with topk, per-type
 * partition) and the C++ build_nlist_gpu / format_nbor_list_gpu
 * (source/lib/include/neighbor_list.h, fmt_nlist.h:21-30).
 * ------------------------------------------------------------------------ */

/* One frame of a (possibly ragged) batch. Filled on the host. */
typedef struct dpk_frame {
  double box[9];        /* rows are the cell vectors (pbc) */
  double rec[9];        /* inverse of box (pbc) */
  double origin[3];     /* binning origin (non-pbc) */
  double inv_extent[3]; /* 1 / binning extent per axis (non-pbc) */
  int atom_off;         /* first atom of the frame in the packed arrays */
  int nloc;             /* atoms in the frame */
  int cell_off;         /* first cell of the frame in the packed cell arrays */
  int nc[3];            /* cells per axis (fractional grid) */
  int rng[3];           /* cells searched on each side per axis */
  int pbc;              /* 1: periodic, wrap + images; 0: open boundaries */
  int img_off;          /* offset into the per-frame image tables */
  int nb[3];            /* image shift range per axis: [-nb, nb] */
  int ns;               /* number of images (shift list length) */
  int out_off;          /* first output row (atom) -- equals atom_off */
} dpk_frame;

/* Bytes of scratch the search needs for `natoms` atoms and `ncells` cells. */
size_t dpk_nlist_workspace_bytes(int natoms, int ncells);

/*
 * Build the formatted neighbour list of every atom of a packed batch.
 *   coord [natoms*3]      input coordinates (unwrapped)
 *   atype [natoms]        types, < 0 = virtual atom (never a neighbour, row -1)
 *   frame_of [natoms]     frame index of each atom
 *   frames [nframes]      dpk_frame descriptors (device copy)
 *   img_pos [..]          per frame, at img_off: table of (2nb+1)^3 ints mapping
 *                         shift (sx,sy,sz) -> image position s
 *   shift_vec [..*3]      per frame, at img_off*3: shift_idx[s] @ box, FP64
 *   sel_host [ntypes]     per-type slot counts; nnei = sum(sel)
 *   max_cand              per-atom candidate capacity (64..4096)
 * Outputs: nlist [natoms*nnei] extended indices (frame-local), wrapped
 * coordinates coord_wrapped [natoms*3]; status_dev[0] receives the largest
 * within-rcut neighbour count seen (check it against max_cand on the host).
 * Order inside each type block is ascending FP64 distance
 * (|d| = sqrt(fma(z,z,fma(y,y,x*x)))), ties broken by extended index; the
 * nnei nearest over all types are kept before the per-type truncation.
 */
int dpk_nlist_build(int* nlist, double* coord_wrapped, int* status_dev,
                    const double* coord, const int* atype, const int* frame_of,
                    const dpk_frame* frames, int nframes, int natoms, int ncells,
                    const int* img_pos, const double* shift_vec, double rcut,
                    const int* sel_host, int ntypes, int max_cand, void* workspace,
                    size_t workspace_bytes, dpk_stream_t stream);

/*
 * dpk_nlist_build without the workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The cell-list workspace lives in a 32 MiB static device array (one per
 * GPU, held until the process ends), so dpk_nlist_workspace_bytes(natoms,
 * ncells) must not exceed 32 MiB (about 2 million atoms); larger systems are
 * rejected with DPK_ERR_ARG.  All calls on one GPU share that array: calls
 * that overlap on one GPU (different streams or threads) race.  Calls no
 * allocating routine; the first call on a device looks up the array's
 * address.  Same outputs as dpk_nlist_build.
 */
int dpk_nlist_build_noalloc(int* nlist, double* coord_wrapped, int* status_dev,
                            const double* coord, const int* atype, const int* frame_of,
                            const dpk_frame* frames, int nframes, int natoms, int ncells,
                            const int* img_pos, const double* shift_vec, double rcut,
                            const int* sel_host, int ntypes, int max_cand,
                            dpk_stream_t stream);

/*
 * Format a caller's neighbour list (forward_lower input: an MD engine's list,
 * built with rcut + skin, any order) as the model needs it: replaces
 * deepmd/pt/model/model/make_model.py:541-658 (format_nlist / _format_nlist)
 * and nlist.py:304-337 (nlist_distinguish_types).
 *   nlist_in [nframes*nloc, nnei_in]  extended indices into the frame's
 *                         ext_coord [nall*3] / ext_atype [nall], -1 = empty
 *   extra_sort            force the distance sort (the sort also runs when
 *                         nnei_in > sum(sel))
 * Entries with |d| > rcut are dropped (FP64 norm, torch CPU rounding).  With
 * the sort: ascending (distance, index), the sum(sel) nearest kept.  Without:
 * the input order is kept.  Then each type block takes its type's entries in
 * that order, cut at sel[t]; entries of type < 0 fill no block.  Output
 * nlist [nframes*nloc, sum(sel)].  status_dev[0] = largest within-rcut count,
 * status_dev[1] = 1 if an index >= nall was seen (such entries are dropped;
 * the reference's gather would fail on them).  nnei_in <= 2048.
 */
int dpk_nlist_format(int* nlist, int* status_dev, const int* nlist_in, int nnei_in,
                     const double* ext_coord, const int* ext_atype, int nframes,
                     int nloc, int nall, double rcut, const int* sel_host,
                     int ntypes, int extra_sort, dpk_stream_t stream);

/*
 * Model-level pair exclusion (pair_exclude_types of the atomic model):
 * replaces deepmd/pt/model/atomic_model/base_atomic_model.py:380-383
 * (nlist = where(PairExcludeMask(nlist, extended_atype) == 1, nlist, -1),
 * exclude_mask.py:64-140).  Every entry of nlist [natoms, nnei] whose
 * (centre type, neighbour type) pair has pair_mask[(ntypes+1)^2] == 0
 * becomes -1 in place; the neighbour of entry e of atom i is atom
 * frames[f].atom_off + e % frames[f].nloc (f = frame_of[i]), a centre of
 * type < 0 uses row ntypes.  Run it on the formatted list.
 */
int dpk_nlist_pair_exclude(int* nlist, const int* atype, const int* frame_of,
                           const dpk_frame* frames, int natoms, int nnei,
                           const int* pair_mask, int ntypes, dpk_stream_t stream);

/* ------------------------------------------------------------------------ *
 * prod_env_mat_a (op form, explicit extended coordinates).
 * Replaces deepmd::prod_env_mat_a_gpu (source/lib/include/prod_env_mat.h:92-112)
 * given an already formatted nlist, with the normalisation of
 * source/lib/src/prod_env_mat.cc:95-119 and the PT padding rule
 * (deepmd/pt/model/descriptor/env_mat.py:33-55,116-118: padded slots give
 * em = -avg/std, em_deriv = 0, rij = 0).
 *   coord [nf, nall*3], atype [nf, nall] (centre types = first nloc),
 *   avg/std [ntypes, nnei*4].  A centre of type < 0 or >= ntypes gets zero
 *   rows; neighbour indices outside [0, nall) are treated as empty slots.
 *   protection >= 0 (the descriptor's env_protection, env_mat.py:38-45):
 *   with u = r + protection the row is sw (1/u, x/u^2, y/u^2, z/u^2); 0 is
 *   the C++ op's form.
 * ------------------------------------------------------------------------ */
int dpk_prod_env_mat_a_f64(double* em, double* em_deriv, double* rij,
                           const double* coord, const int* atype, const int* nlist,
                           const double* avg, const double* std, int nframes,
                           int nloc, int nall, int nnei, int ntypes, double rcut,
                           double rcut_smth, double protection, dpk_stream_t stream);
int dpk_prod_env_mat_a_f32(float* em, float* em_deriv, float* rij,
                           const float* coord, const int* atype, const int* nlist,
                           const float* avg, const float* std, int nframes,
                           int nloc, int nall, int nnei, int ntypes, double rcut,
                           double rcut_smth, double protection, dpk_stream_t stream);

/* ------------------------------------------------------------------------ *
 * prod_force_se_a / prod_virial_se_a and their gradients.
 * Replace deepmd::prod_force_a_gpu (source/lib/include/prod_force.h:72-79),
 * prod_virial_a_gpu (prod_virial.h:27-36), prod_force_grad_a_gpu
 * (prod_force_grad.h:25-32) and prod_virial_grad_a_gpu (prod_virial_grad.h:25-32).
 *   net_deriv [nf, nloc*nnei*4], in_deriv [nf, nloc*nnei*12],
 *   rij [nf, nloc*nnei*3], nlist [nf, nloc*nnei].
 * prod_force: force [nf, nall*3] (overwritten).  Neighbour indices outside
 *   [0, nall) add nothing (no device-side fault, no host check needed).
 * prod_virial: virial [nf, 9], atom_virial [nf, nall*9] (overwritten;
 *   NULL: not formed, as a training step that needs only the virial);
 *   workspace >= dpk_prod_virial_workspace_bytes(nf, nloc).
 * force_grad: grad [nf, ngrad*3] -> grad_net [nf, nloc*nnei*4]; with
 *   ngrad == nloc a neighbour index >= nloc is taken modulo nloc (the
 *   reference's prod_force_grad.cc:54-57), with ngrad == nall it indexes the
 *   extended rows directly (exact adjoint of prod_force for extended forces).
 * virial_grad: grad [nf, 9] -> grad_net [nf, nloc*nnei*4].
 * ------------------------------------------------------------------------ */
int dpk_prod_force_a_f64(double* force, const double* net_deriv,
                         const double* in_deriv, const int* nlist, int nframes,
                         int nloc, int nall, int nnei, dpk_stream_t stream);
size_t dpk_prod_virial_workspace_bytes(int nframes, int nloc);
int dpk_prod_virial_a_f64(double* virial, double* atom_virial,
                          const double* net_deriv, const double* in_deriv,
                          const double* rij, const int* nlist, int nframes,
                          int nloc, int nall, int nnei, void* workspace,
                          size_t workspace_bytes, dpk_stream_t stream);
int dpk_prod_force_grad_a_f64(double* grad_net, const double* grad,
                              const double* in_deriv, const int* nlist,
                              int nframes, int nloc, int nnei, int ngrad,
                              dpk_stream_t stream);
int dpk_prod_virial_grad_a_f64(double* grad_net, const double* grad,
                               const double* in_deriv, const double* rij,
                               const int* nlist, int nframes, int nloc,
                               int nnei, dpk_stream_t stream);

/* ------------------------------------------------------------------------ *
 * tabulate_fusion_se_a forward / grad / grad_grad (is_sorted padding fold).
 * Replace deepmd::tabulate_fusion_se_a_gpu, _grad_gpu, _grad_grad_gpu
 * (source/lib/include/tabulate.h:100-140; CPU semantics tabulate.cc:162-447).
 *   table [nspline, 6*L] (interval indices are clamped to [0, nspline)),
 *   table_info_host [6] (HOST: lower, upper, max,
 *   stride0, stride1, check_freq), em_x [nloc, nnei], em [nloc, nnei, nd],
 *   out / dz_dy [nloc, nd, L], dy [nloc, nd, L]. nd in {4, 9, 16, 25}.
 * ------------------------------------------------------------------------ */
int dpk_tabulate_fusion_se_a_f64(double* out, const double* table, int nspline,
                                 const double* table_info_host, const double* em_x,
                                 const double* em, int nloc, int nnei, int L,
                                 int nd, dpk_stream_t stream);
int dpk_tabulate_fusion_se_a_f32(float* out, const float* table, int nspline,
                                 const float* table_info_host, const float* em_x,
                                 const float* em, int nloc, int nnei, int L,
                                 int nd, dpk_stream_t stream);
int dpk_tabulate_fusion_se_a_grad_f64(double* dy_dem_x, double* dy_dem,
                                      const double* table, int nspline,
                                      const double* table_info_host,
                                      const double* em_x, const double* em,
                                      const double* dy, int nloc, int nnei, int L,
                                      int nd, dpk_stream_t stream);
int dpk_tabulate_fusion_se_a_grad_f32(float* dy_dem_x, float* dy_dem,
                                      const float* table, int nspline,
                                      const float* table_info_host,
                                      const float* em_x, const float* em,
                                      const float* dy, int nloc, int nnei, int L,
                                      int nd, dpk_stream_t stream);
int dpk_tabulate_fusion_se_a_grad_grad_f64(double* dz_dy, const double* table, int nspline,
                                           const double* table_info_host,
                                           const double* em_x, const double* em,
                                           const double* dz_dy_dem_x,
                                           const double* dz_dy_dem, int nloc,
                                           int nnei, int L, int nd,
                                           dpk_stream_t stream);
int dpk_tabulate_fusion_se_a_grad_grad_f32(float* dz_dy, const float* table, int nspline,
                                           const float* table_info_host,
                                           const float* em_x, const float* em,
                                           const float* dz_dy_dem_x,
                                           const float* dz_dy_dem, int nloc,
                                           int nnei, int L, int nd,
                                           dpk_stream_t stream);

/* ------------------------------------------------------------------------ *
 * Fused se_e2_a descriptor path (DeepPot.eval / forward_lower hot path).
 * Replaces the eager chain of deepmd/pt/model/descriptor/env_mat.py:18-119,
 * se_a.py:742-858 (embedding nets, R^T G, G^T R R^T G) and its autograd
 * backward (transform_output.py:68-99) for one packed batch.
 *
 * Compute precision `prec`: DPK_PREC_F64 (FP64 everywhere), DPK_PREC_F32
 * (FP32 MFMA, FP32 accumulation), DPK_PREC_BF16 (BF16 MFMA operands, FP32
 * accumulation and activations), DPK_PREC_F16 (the same with FP16 MFMA
 * operands: 11-bit significand, same MFMA rate; the embedding backward
 * scales each atom's gradient operands by a power of two into FP16's normal
 * range), DPK_PREC_F16X2 (FP16 weights; every data operand -- activations,
 * tangents, env-matrix rows, G, gradients -- split into hi = fp16(x) and
 * lo = fp16(x - hi), two or three MFMAs per product, ~22-bit operands).
 * Geometry (distances, switch, derivatives) is FP64 in all.
 * "real" below is double for F64 and float otherwise; "16-bit" is BF16 or
 * FP16 bits as `prec` says (F16X2 tensors exchanged with the caller are
 * described where they differ).
 * ------------------------------------------------------------------------ */
enum { DPK_PREC_F64 = 0, DPK_PREC_F32 = 1, DPK_PREC_BF16 = 2, DPK_PREC_F16 = 3,
       DPK_PREC_F16X2 = 4 };

/*
 * Env matrix of the fused path with implicit images: neighbour e of an atom
 * in frame f sits at coord_wrapped[atom_off + e % nloc] + shift_vec[(img_off +
 * e / nloc)*3 ..]. Writes em [natoms, nnei, 4] (real) normalised by
 * avg/std [ntypes, nnei*4] (FP64) of the centre type.
 * pair_mask [(ntypes+1)^2] (int, NULL = no exclusion) is the type-pair table
 * of deepmd/pt/utils/exclude_mask.py:64-140 (PairExcludeMask): a neighbour j
 * of centre i with pair_mask[type_i*(ntypes+1) + type_j] == 0 gets a zero row
 * (the rr * mask of se_a.py:811-819); virtual centres (atype -1) use row
 * ntypes, padding slots are never masked.  protection as for
 * dpk_prod_env_mat_a_f64.
 */
int dpk_env_mat_fused(void* em, int prec, const double* coord_wrapped,
                      const int* atype, const int* frame_of,
                      const dpk_frame* frames, const int* nlist,
                      const double* shift_vec, const double* avg,
                      const double* std, int natoms, int nnei, double rcut,
                      double rcut_smth, double protection,
                      const int* pair_mask, int ntypes, dpk_stream_t stream);

/*
 * Activation functions of the embedding and fitting nets
 * (deepmd/pt/utils/utils.py:174-212): tanh everywhere; the others in the
 * FP32 / FP64 precisions (the 16-bit kernels fold tanh into their weight
 * images and reject them).  act_param: the silut / custom_silu threshold.
 */
enum {
  DPK_ACT_TANH = 0,
  DPK_ACT_LINEAR = 1,   /* "linear" / "none" */
  DPK_ACT_RELU = 2,
  DPK_ACT_RELU6 = 3,
  DPK_ACT_SOFTPLUS = 4,
  DPK_ACT_SIGMOID = 5,
  DPK_ACT_SILU = 6,
  DPK_ACT_GELU = 7,     /* "gelu" / "gelu_tf": tanh approximation */
  DPK_ACT_SILUT = 8     /* "silut" / "custom_silu[:threshold]" */
};

/*
 * Packed embedding parameters (built by the host; one block per net):
 *   net k at offset k * dpk_embed_net_elems(P1,P2,P3) elements of "real":
 *   w1[P1] b1[P1] | w2t[P2][P1] b2[P2] idt2[P2] | w3t[P3][P2] b3[P3] idt3[P3]
 * where P1/P2/P3 are the internal (16-padded) widths (only 32/64/128
 * supported), wNt is the transposed weight (out x in) in the internal channel
 * order, padded channels have zero weight and bias, idt is 1 when the layer
 * has none, and the resnet rule per layer is given by res2/res3 (0 none,
 * 2 y += (x, x); 1, y += x, is rejected: it is not representable in the
 * fixed internal widths).  has_idt = 0 promises idt == 1 for both layers
 * (no time step); the BF16 kernels then skip the multiply.  act / act_param
 * (DPK_ACT_*, silut threshold): the activation of every layer (tanh only in
 * the 16-bit precisions).
 * BF16 reads its parameters from a device "net image" instead of `params`:
 * dpk_embed_image_bytes(P1,P2,P3) bytes per net, written by
 * dpk_embed_pack_bf16 from the FP32 packed parameters (FP32 vectors, BF16
 * transposed weights in MFMA K order, rows padded for LDS).  `params` is then
 * unused (may be NULL); for F32/F64 `image` is unused.
 * ch_out [P3] maps internal output channel -> true channel (-1 = padding).
 * net_of[ntypes_c * ntypes]: net index (< nnets) for (centre type,
 * neighbour-type block); nnets = ntypes (type_one_side) or ntypes^2.
 * sec_host[ntypes+1]: type block starts.
 */
size_t dpk_embed_net_elems(int P1, int P2, int P3);
/* Bytes of one BF16 net image (0 if the widths are unsupported).  Its
 * weights and biases are scaled by 2 log2(e), so that an MFMA started from
 * the scaled bias yields the tanh argument directly. */
size_t dpk_embed_image_bytes(int P1, int P2, int P3);
/* image [nnets * dpk_embed_image_bytes] from params (FP32 packed layout). */
int dpk_embed_pack_bf16(void* image, const float* params, int nnets, int P1,
                        int P2, int P3, dpk_stream_t stream);
/* The same image with FP16 weights, for prec = DPK_PREC_F16 and F16X2. */
int dpk_embed_pack_f16(void* image, const float* params, int nnets, int P1,
                       int P2, int P3, dpk_stream_t stream);

/* Forward: xyz [natoms, 4, d3] (real) = (1/nnei) sum_slots R^T G.
 * nlist (optional, BF16 only): the formatted neighbour list [natoms, nnei];
 * each type block is then computed up to its last real neighbour, and the
 * slots after it (index -1, env-matrix rows -avg/std: constants per centre
 * type and slot, as the reference's padding fold in tabulate.cc:162-341
 * exploits) are taken from
 * pad (optional, needs nlist; NULL when every padding row is zero):
 * [ntypes, nnei, 4, 128] FP32, pad[ti][jj][c][k] = sum over the slots
 * jj..end of jj's block of R[c] G[k] (padding rows of centre type ti,
 * internal channel k of the packed net image, before the 1/nnei). */
int dpk_embed_fwd(void* xyz, int prec, const void* em, const int* atype,
                  const void* params, const void* image, const int* net_of,
                  const int* ch_out, int ntypes, int nnets, int P1, int P2,
                  int P3, int res2, int res3, int has_idt, int act,
                  double act_param, int d3, const int* sec_host, int natoms,
                  int nnei, const int* nlist, const float* pad,
                  dpk_stream_t stream);

/*
 * Backward: given gxyz = dE/dxyz [natoms, 4, d3] (real), writes
 * net_deriv = dE/dem [natoms, nnei, 4] (real), recomputing the embedding and
 * its tangent dG/ds (activations are never stored).  nlist (optional, BF16
 * only) as for dpk_embed_fwd: the slots after a block's last real neighbour
 * get net_deriv 0 (no force or virial term reads them).
 */
int dpk_embed_bwd(void* net_deriv, int prec, const void* em, const void* gxyz,
                  const int* atype, const void* params, const void* image,
                  const int* net_of, const int* ch_out, int ntypes, int nnets,
                  int P1, int P2, int P3, int res2, int res3, int has_idt,
                  int act, double act_param, int d3, const int* sec_host,
                  int natoms, int nnei, const int* nlist, dpk_stream_t stream);

/*
 * Training: the backward passes of the two embedding functions a force-loss
 * training step differentiates (replacing torch's double backward through
 * DescrptBlockSeA.forward, deepmd/pt/model/descriptor/se_a.py:784-858, with
 * create_graph=True as in deepmd/pt/model/model/make_model.py:461 and
 * transform_output.py:74-86):
 *   X(params)     : xyz = (1/nnei) sum_j R_j^T G_j          (dpk_embed_fwd)
 *   N(params, gx) : net_deriv = dE/dem given gx = dE/dxyz    (dpk_embed_bwd)
 * BF16 operands, FP32 accumulation, internal widths (32, 64, 128), nets
 * without time steps (has_idt = 0, else DPK_ERR_UNSUPPORTED).
 *
 * Backward image of the nets (the weights W2, W3 in the order their
 * transposed products read them): dpk_embed_train_bimage_bytes() bytes per
 * net, written from the FP32 packed parameters by dpk_embed_train_pack_bwd.
 */
size_t dpk_embed_train_bimage_bytes(void);
int dpk_embed_train_pack_bwd(void* bimage, const float* params, int nnets,
                             dpk_stream_t stream);
/* FP32 elements of the partial-gradient workspace for nnets x nx workgroups. */
size_t dpk_embed_train_parts_elems(int nnets, int nx);
/*
 * tangent = 0 (backward of X): gmat = dL/dxyz [natoms, 4, d3]; writes
 *   gparams = dL/dparams [nnets, dpk_embed_net_elems] (packed layout, FP32;
 *   the idt entries are 0).
 * tangent = 1 (backward of N): gmat = gx [natoms, 4, d3] (the N input),
 *   vnd = dL/dnet_deriv [natoms, nnei, 4]; writes gparams and
 *   ggx [nblocks, natoms, 4, d3] = dL/dgx per type block (entry (b, i) is
 *   written when atom i belongs to a net of block b; sum over b).
 * em [natoms, nnei, 4] FP32; image / bimage: the nets' forward image
 * (dpk_embed_pack_bf16) and backward image.  Net k serves the atoms
 * atoms[net_atom_off[k] .. net_atom_off[k+1]) over the slots
 * [net_b0[k], net_b1[k]) of type block net_blk[k] (host arrays, nnets <= 16;
 * type_one_side: net k = block k, every atom).  parts: workspace of
 * dpk_embed_train_parts_elems(nnets, nx) FP32 (nx workgroups per net; the
 * partials are summed in a fixed order, so results are reproducible).
 */
int dpk_embed_train_vjp(float* gparams, float* ggx, float* parts, int tangent,
                        const float* em, const float* gmat, const float* vnd,
                        const void* image, const void* bimage, const int* atoms,
                        const int* net_atom_off, const int* net_b0,
                        const int* net_b1, const int* net_blk,
                        const int* ch_out, int nnets, int nx, int res2,
                        int res3, int has_idt, int d3, int natoms, int nnei,
                        dpk_stream_t stream);

/*
 * Compressed embedding in the fused path (se_a.py:795-843 with
 * tabulate_fusion_se_a per type block, tabulate.cc:162-341 semantics incl.
 * the padding fold per block): xyz [natoms, 4, L] = (1/nnei) sum over type
 * blocks of the block's table applied to em_x = em[..., 0] and em; the
 * backward writes net_deriv = dE/dem [natoms, nnei, 4] from gxyz = dE/dxyz.
 * prec: DPK_PREC_F64 (double tables/em/xyz) or DPK_PREC_F32 (float).
 * tables: all nets' tables [nspline_k, 6 L] concatenated (device, element
 * offsets table_off[k]); nspline[k], table_off[k] and info[6 k .. 6 k + 5] =
 * (lower, upper, upper*extrapolate, stride0, stride1, check_frequency) are
 * host arrays; nnets <= 16.  net_of / sec_host as for dpk_embed_fwd.
 */
int dpk_tab_embed_fwd(void* xyz, int prec, const void* em, const int* atype,
                      const void* tables, const long long* table_off,
                      const int* nspline, const double* info, const int* net_of,
                      int ntypes, int nnets, const int* sec_host, int L,
                      int natoms, int nnei, dpk_stream_t stream);
int dpk_tab_embed_bwd(void* net_deriv, int prec, const void* em,
                      const void* gxyz, const int* atype, const void* tables,
                      const long long* table_off, const int* nspline,
                      const double* info, const int* net_of, int ntypes,
                      int nnets, const int* sec_host, int L, int natoms,
                      int nnei, dpk_stream_t stream);

/* Descriptor D [natoms, d3, m2] = xyz^T xyz[:, :, :m2] (out dtype = `out_prec`). */
int dpk_descriptor_fwd(void* desc, int out_prec, const void* xyz, int prec,
                       int natoms, int d3, int m2, dpk_stream_t stream);
/* gxyz [natoms, 4, d3] (real) from gdesc = dE/dD [natoms, d3, m2] (`g_prec`). */
int dpk_descriptor_bwd(void* gxyz, int prec, const void* gdesc, int g_prec,
                       const void* xyz, int natoms, int d3, int m2,
                       dpk_stream_t stream);

/* ------------------------------------------------------------------------ *
 * Energy fitting around hipBLASLt GEMMs (fitting.py:734-899, mlp.py:188-219,
 * se_a.py:844-850).  Rows are "slots": real atoms grouped by type, slot i
 * being atom order[i], so that each type's layer is one contiguous GEMM
 * z = x W (done by the caller).  Precision `prec` as above: the GEMM type
 * "gemm" is bf16 / fp16 / float / double for BF16 / F16 / F32 / F64; "real"
 * is float for BF16, F16 and F32, double for F64.  res: 0 none, 1 y += x,
 * 2 y += (x, x).  DPK_PREC_F16X2 is accepted by the two descriptor entries
 * only (its layers run in dpk_fit_tail_f16x2).
 * ------------------------------------------------------------------------ */
/* desc [nslots, d3*m2] (gemm) = xyz^T xyz[:, :, :m2] of atom order[slot]
 * (xyz [natoms, 4, d3], real); d3 <= 256, m2 <= 16.  F16X2: desc is
 * [nslots, 2*d3*m2] FP16, each row the hi parts fp16(D) then the lo parts
 * fp16(D - hi) (m2 = 8 or 16); the descriptor backward then takes an FP32
 * gdesc [nslots, d3*m2]. */
int dpk_fit_descriptor_fwd(void* desc, int prec, const void* xyz,
                           const int* order, int nslots, int d3, int m2,
                           dpk_stream_t stream);
/* gxyz [natoms, 4, d3] (real) = dE/dxyz of atom order[slot] from
 * gdesc = dE/ddesc [nslots, d3*m2] (gemm); other atoms untouched. */
int dpk_fit_descriptor_bwd(void* gxyz, int prec, const void* gdesc,
                           const void* xyz, const int* order, int nslots,
                           int d3, int m2, dpk_stream_t stream);
/* Hidden layer epilogue on n rows of width h: y = act(z + bias) * idt +
 * res(y_prev [n, hin]) (real), yg = y (gemm; may be NULL when gemm == real),
 * sd = idt act'(z + bias) (gemm).  idt NULL means 1.  act / act_param:
 * DPK_ACT_* and the silut threshold. */
int dpk_fit_layer_fwd(void* y, void* yg, void* sd, int prec, const void* z,
                      const void* bias, const void* idt, const void* y_prev,
                      int n, int h, int hin, int res, int act, double act_param,
                      dpk_stream_t stream);
/* Output layer: atom_energy[order[i]] = y_i . wout + ebias (FP64 out; ebias =
 * output bias + bias_atom_e + out_bias of the type), gz = wout * sd (gemm). */
int dpk_fit_energy(double* atom_energy, void* gz, int prec, const void* y,
                   const void* sd, const void* wout, double ebias,
                   const int* order, int n, int h, dpk_stream_t stream);
/* Backward epilogue of a hidden layer of width h: gy = mm + res_next^T(g),
 * where mm = g_z(next) W(next)^T (gemm) and g is the next layer's output
 * gradient gy_next [n, hnext] (real) or, for the last hidden layer, the
 * broadcast vector wvec [hnext] (= wout); gz = gy * sd (gemm).  gy or gz may
 * be NULL when not needed. */
int dpk_fit_layer_bwd(void* gy, void* gz, int prec, const void* mm,
                      const void* gy_next, const void* wvec, const void* sd,
                      int n, int h, int hnext, int res_next,
                      dpk_stream_t stream);
/* Fused BF16 fitting tail of one type segment, replacing, for nets of L = 2
 * or 3 hidden layers of width 256, the chain dpk_fit_layer_fwd (layer 1..L),
 * the layer 2..L GEMMs, dpk_fit_energy, the backward GEMMs of layers L..2
 * and dpk_fit_layer_bwd (the same math: GeneralFitting._forward_common,
 * deepmd/pt/model/task/fitting.py:734-899, and its autograd).
 *   z1 [n, 256]     BF16 layer-1 pre-activations (z1 = D W_1, bias not added)
 *   gz1 [n, 256]    BF16 out: dE/dz1 (the caller forms dE/dD = gz1 W_1^T)
 *   atom_energy     FP64 out at order[i] for slot i: y_L . w_out + ebias
 *   net             device image of dpk_fit_tail_bytes(L), 16-byte aligned:
 *                   for l = 2..L the 256 x 256 matrix W_l^T, then for
 *                   l = 2..L the matrix W_l, each as BF16 MFMA fragments
 *                   (entry ((mb*8 + kc)*64 + 16*g + m)*8 + j holds row
 *                   16*mb + m, column 32*kc + 16*(j/4) + 4*g + j%4 of the
 *                   matrix), then FP32 vectors of 256: b_1, idt_1, then
 *                   (b_l, idt_l) for l = 2..L, then w_out (idt = 1 when the
 *                   layer has none; padded channels zero).
 *   res_mask        bit l-2 set: layer l adds y_{l-1} (same width resnet).
 * Needs 160 KB of LDS per workgroup; no counterpart entry in the reference
 * (its fitting runs as separate torch ops). */
size_t dpk_fit_tail_bytes(int L);
int dpk_fit_tail_bf16(double* atom_energy, void* gz1, const void* z1,
                      const void* net, int L, int res_mask, double ebias,
                      const int* order, int n, dpk_stream_t stream);
/* The same with FP16 z1, gz1 and matrix fragments (prec = DPK_PREC_F16). */
int dpk_fit_tail_f16(double* atom_energy, void* gz1, const void* z1,
                     const void* net, int L, int res_mask, double ebias,
                     const int* order, int n, dpk_stream_t stream);
/* Split mode (DPK_PREC_F16X2, the FP16 image): z1 [n, 256] is FP32 (16-byte
 * aligned) and gz1 [n, 512] FP16 holds per row the hi parts fp16(g) then the
 * lo parts fp16(g - hi), so that [gz1_hi | gz1_lo] x [W_1^T; W_1^T] (K = 512)
 * forms dE/dD to ~22 bits. */
int dpk_fit_tail_f16x2(double* atom_energy, void* gz1, const void* z1,
                       const void* net, int L, int res_mask, double ebias,
                       const int* order, int n, dpk_stream_t stream);

/*
 * Split-FP16 fitting layer 1 fused with the descriptor (prec DPK_PREC_F16X2):
 * replaces dpk_fit_descriptor_fwd + the layer-1 GEMM and the dE/dD GEMM +
 * dpk_fit_descriptor_bwd, without writing D or dE/dD to memory (the same
 * math: se_a.py:844-850, fitting.py:734-899 and their autograd).  For one
 * type segment of n slots (slot i = atom order[i]); m2 = 16, width h = 256
 * (a padded fitting width), 16 <= d3 <= 256.
 *   image [dpk_fit_l1_image_bytes(d3, h)], 16-byte aligned: FP16 MFMA
 *         fragments of W1, written by dpk_fit_l1_pack from the FP32 matrix
 *         w1 [16*d3, h] (row q = 16 k + m of D, column h).
 *   fwd:  z1 [n, 256] FP32 = D W1 (no bias: the fitting tail adds it).
 *   bwd:  gxyz [natoms, 4, d3] FP32 rows of the segment's atoms = dE/dxyz
 *         from gz1 [n, 512] (hi | lo FP16 halves, dpk_fit_tail_f16x2).
 * No counterpart entry in the reference (its fitting runs as torch ops).
 */
size_t dpk_fit_l1_image_bytes(int d3, int h);
int dpk_fit_l1_pack(void* image, int prec, const float* w1, int d3, int m2, int h,
                    dpk_stream_t stream);
int dpk_fit_l1_fwd(float* z1, int prec, const float* xyz, const int* order, int n, int d3,
                   int m2, const void* image, int h, dpk_stream_t stream);
int dpk_fit_l1_bwd(float* gxyz, int prec, const void* gz1, const float* xyz, const int* order,
                   int n, int d3, int m2, const void* image, int h, dpk_stream_t stream);

/*
 * Forces and virial from net_deriv with recomputed FP64 geometry (implicit
 * images): force [natoms*3] (FP64, accumulated: caller zeroes it),
 * per-atom virial rows (FP64, 9 per centre atom) written to `workspace`,
 * reduced per frame by dpk_frame_sum (ncomp 9); atom_virial [natoms*9]
 * optional (NULL to skip), accumulated on the owning local atom of each
 * neighbour (the reference's extended atom virial after ghost -> local).
 * pair_mask / ntypes / protection as in dpk_env_mat_fused: excluded slots (a
 * constant zero env-matrix row) contribute nothing.
 */
size_t dpk_force_virial_workspace_bytes(int natoms);
int dpk_force_virial_fused(double* force, double* atom_virial, const void* net_deriv,
                           int prec, const double* coord_wrapped,
                           const int* atype, const int* frame_of,
                           const dpk_frame* frames, const int* nlist,
                           const double* shift_vec, const double* std,
                           int natoms, int nnei, double rcut, double rcut_smth,
                           double protection, void* workspace,
                           size_t workspace_bytes, const int* pair_mask,
                           int ntypes, dpk_stream_t stream);
/*
 * Forces and virial of an engine's extended region (the MD-engine path:
 * explicit ghosts, nlist [natoms, nnei] of extended-atom indices in
 * [0, nall), coord = the extended coordinates [nall, 3]) with the
 * neighbour sums gathered, not added with FP64 atomics: per call the slots
 * are binned by neighbour into fixed-capacity buckets (slots beyond a full
 * bucket are added with atomics at the end), a first pass writes each slot's
 * neighbour share and the centre's own force, a second pass sums every
 * extended atom's bucket.
 * force [nall, 3] and atom_virial [nall, 9] (optional, NULL to skip; the
 * neighbour rows, as dpk_force_virial_fused) are overwritten; vrow
 * [natoms, 9] receives the per-centre virial rows (reduce with
 * dpk_frame_sum).  pair_mask / ntypes / protection as in dpk_env_mat_fused.
 * No counterpart in the reference's kernels (its prod_force adds with
 * atomics); the same sums as dpk_force_virial_fused in another order.
 */
size_t dpk_force_virial_gather_workspace_bytes(int natoms, int nnei, int nall);
int dpk_force_virial_gather(double* force, double* atom_virial, double* vrow,
                            const void* net_deriv, int prec, const double* coord,
                            const int* atype, const int* nlist, const double* std,
                            int natoms, int nnei, int nall, double rcut,
                            double rcut_smth, double protection,
                            const int* pair_mask, int ntypes, void* workspace,
                            size_t workspace_bytes, dpk_stream_t stream);
/*
 * Same result as dpk_force_virial_fused, organised by the neighbour-list
 * cells: one workgroup per cell of the grid built by the last
 * dpk_nlist_build (its workspace is passed back), neighbour forces summed in
 * LDS per 3x3x3 cell block and added to `force` once per workgroup.
 * img_shift [total images, 3] holds the integer shift of each image (the
 * rows of shift_vec in lattice units).  force accumulates (caller zeroes).
 */
int dpk_force_virial_cells(double* force, double* atom_virial, const void* net_deriv,
                           int prec, const double* coord_wrapped,
                           const int* atype, const int* frame_of,
                           const dpk_frame* frames, const int* nlist,
                           const double* shift_vec, const int* img_shift,
                           const double* std, int natoms, int ncells, int nnei,
                           double rcut, double rcut_smth, double protection,
                           void* nlist_workspace,
                           size_t nlist_workspace_bytes, void* workspace,
                           size_t workspace_bytes, const int* pair_mask,
                           int ntypes, dpk_stream_t stream);

/*
 * out [nframes, ncomp] = per-frame sum of in [natoms, ncomp] (FP64) over the
 * frame's atom range (frames[f].atom_off, .nloc), in a fixed order (two
 * passes through `workspace`; no atomics, so results are reproducible).
 * ncomp is 1 (atom energies -> frame energies, the torch.sum over atoms of
 * deepmd/pt/model/model/transform_output.py:176-178) or 9 (per-atom virial
 * rows -> frame virial). max_nloc >= every frames[f].nloc.
 */
size_t dpk_frame_sum_workspace_bytes(int nframes, int max_nloc, int ncomp);
int dpk_frame_sum(double* out, const double* in, int ncomp,
                  const dpk_frame* frames, int nframes, int max_nloc,
                  void* workspace, size_t workspace_bytes, dpk_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DPK_H */
