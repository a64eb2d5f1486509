// libdpk_deeppot: a se_e2_a model driven from C++ through libdpk's C ABI, as
// an MD engine calls deepmd::DeepPotPT::compute with its own neighbour list
// (source/api_cc/src/DeepPotPT.cc:171-378).  Per step:
//   dpk_nlist_format      the engine's list -> sorted, rcut-cut, per-type rows
//   dpk_env_mat_fused     environment matrix (FP64 geometry)
//   dpk_embed_fwd         embedding + R^T G            (split FP16)
//   dpk_fit_l1_fwd        descriptor + fitting layer 1 (per type segment)
//   dpk_fit_tail_f16x2    layers 2..L, energy, backward to dE/dz1
//   dpk_fit_l1_bwd        dE/dz1 -> dE/dxyz
//   dpk_embed_bwd         dE/dxyz -> dE/d em
//   dpk_force_virial_gather, dpk_frame_sum
// The region is "extended": an index frame with nloc = nall atoms and one
// image, so a neighbour index is the engine's atom index and forces land on
// local and ghost atoms alike.
#include <hip/hip_runtime_api.h>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "dpk.h"
#include "dpk_deeppot.h"

namespace {

thread_local char g_err[1024] = {0};

void set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

struct DevBuf {
  void* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // (re)allocate to at least `bytes` (contents not kept)
  bool reserve(size_t bytes) {
    if (bytes <= n && p) return true;
    release();
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
      p = nullptr;
      set_err("hipMalloc of %zu bytes failed", bytes);
      return false;
    }
    n = bytes;
    return true;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

}  // namespace

struct dpk_deeppot {
  // model
  int ntypes = 0, nnei = 0, d3 = 0, m2 = 0, prec = 0, fit_prec = 0, nnets = 0;
  int P1 = 0, P2 = 0, P3 = 0, res2 = 0, res3 = 0, has_idt = 0, skip_padding = 0;
  int h1 = 0, layers = 0;
  double rcut = 0, rcut_smth = 0, protection = 0;
  std::vector<int> sel, sec, keep, tail_res;
  std::vector<double> ebias;
  std::map<std::string, DevBuf> blob;  // device copies of the exported arrays
  int device = 0;
  hipStream_t stream = nullptr;
  // system (sized at ago == 0)
  int nloc = -1, nall = -1, nnei_in = -1, nreal = 0;
  std::vector<int> seg_type, seg_start, seg_end;
  DevBuf coord, atype, nlist_in, nlist, status, em, xyz, gxyz, atom_e, nd, force, vrow, virial,
      energy, fsum_ws, frame_of, frames, sums, shift, order, z1, gz, gather_ws;
  bool have_list = false;
};

namespace {

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  set_err("%s: %s", what, hipGetErrorString(e));
  return false;
}

bool dpk_ok(int rc, const char* what) {
  if (rc == DPK_OK) return true;
  set_err("%s: %s", what, dpk_last_error());
  return false;
}

// --- model file ---------------------------------------------------------------
bool read_ints(std::istringstream& ss, std::vector<int>& v) {
  int x;
  v.clear();
  while (ss >> x) v.push_back(x);
  return true;
}

bool load_file(dpk_deeppot* dp, const char* path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    set_err("cannot open %s", path);
    return false;
  }
  std::string line;
  if (!std::getline(f, line) || line != "DPKM 1") {
    set_err("%s: not an exported model (header '%s')", path, line.c_str());
    return false;
  }
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string key;
    ss >> key;
    if (key == "end") return true;
    if (key == "blob") {
      std::string name;
      size_t nb = 0;
      ss >> name >> nb;
      std::vector<char> host(nb);
      if (!f.read(host.data(), (std::streamsize)nb)) {
        set_err("%s: short blob %s", path, name.c_str());
        return false;
      }
      f.get();  // newline after the bytes
      DevBuf& b = dp->blob[name];
      if (!b.reserve(nb)) return false;
      if (!hip_ok(hipMemcpy(b.p, host.data(), nb, hipMemcpyHostToDevice), "upload")) return false;
      continue;
    }
    if (key == "ntypes") ss >> dp->ntypes;
    else if (key == "nnei") ss >> dp->nnei;
    else if (key == "rcut") ss >> dp->rcut;
    else if (key == "rcut_smth") ss >> dp->rcut_smth;
    else if (key == "env_protection") ss >> dp->protection;
    else if (key == "d3") ss >> dp->d3;
    else if (key == "m2") ss >> dp->m2;
    else if (key == "prec") ss >> dp->prec;
    else if (key == "fit_prec") ss >> dp->fit_prec;
    else if (key == "nnets") ss >> dp->nnets;
    else if (key == "P") ss >> dp->P1 >> dp->P2 >> dp->P3;
    else if (key == "res") ss >> dp->res2 >> dp->res3;
    else if (key == "has_idt") ss >> dp->has_idt;
    else if (key == "skip_padding") ss >> dp->skip_padding;
    else if (key == "h1") ss >> dp->h1;
    else if (key == "layers") ss >> dp->layers;
    else if (key == "sel") read_ints(ss, dp->sel);
    else if (key == "sec") read_ints(ss, dp->sec);
    else if (key == "keep") read_ints(ss, dp->keep);
    else if (key == "tail_res") read_ints(ss, dp->tail_res);
    else if (key == "ebias") {
      double x;
      dp->ebias.clear();
      while (ss >> x) dp->ebias.push_back(x);
    } else {
      set_err("%s: unknown key '%s'", path, key.c_str());
      return false;
    }
  }
  set_err("%s: missing 'end'", path);
  return false;
}

bool check_model(dpk_deeppot* dp) {
  if (dp->prec != DPK_PREC_F16X2 || dp->fit_prec != DPK_PREC_F16X2) {
    set_err("the runner takes split-FP16 models (prec %d / %d)", dp->prec, dp->fit_prec);
    return false;
  }
  if (dp->ntypes <= 0 || (int)dp->sel.size() != dp->ntypes ||
      (int)dp->sec.size() != dp->ntypes + 1 || (int)dp->tail_res.size() != dp->ntypes ||
      (int)dp->ebias.size() != dp->ntypes) {
    set_err("inconsistent model header");
    return false;
  }
  const char* need[] = {"davg", "dstd", "emb_params", "emb_image", "net_of", "ch_out"};
  for (const char* n : need)
    if (!dp->blob.count(n)) {
      set_err("model lacks %s", n);
      return false;
    }
  for (int t = 0; t < dp->ntypes; ++t)
    if (!dp->blob.count("l1_" + std::to_string(t)) || !dp->blob.count("tail_" + std::to_string(t))) {
      set_err("model lacks the fitting images of type %d", t);
      return false;
    }
  return true;
}

const void* bp(dpk_deeppot* dp, const std::string& name) {
  auto it = dp->blob.find(name);
  return it == dp->blob.end() ? nullptr : it->second.p;
}

dpk_frame one_frame(int nloc) {
  dpk_frame fr;
  std::memset(&fr, 0, sizeof(fr));
  fr.nloc = nloc;
  fr.ns = 1;
  return fr;
}

// --- per-system buffers and the list, at ago == 0 --------------------------------
bool setup(dpk_deeppot* dp, const int* atype, int nall, int nloc, int inum, const int* ilist,
           const int* numneigh, const int* const* firstneigh) {
  // the engine's list as dense rows of local atoms (InputNlist ->
  // NeighborListData::copy_from_nlist + padding, common.cc:297-379)
  if (inum != nloc) {
    set_err("the list has %d rows for %d local atoms", inum, nloc);
    return false;
  }
  int width = 0;
  for (int ii = 0; ii < inum; ++ii) width = numneigh[ii] > width ? numneigh[ii] : width;
  std::vector<int> rows((size_t)nloc * (width ? width : 1), -1);
  std::vector<char> seen(nloc, 0);
  for (int ii = 0; ii < inum; ++ii) {
    const int i = ilist[ii];
    if (i < 0 || i >= nloc || seen[i]) {
      set_err("ilist[%d] = %d is not a distinct local atom", ii, i);
      return false;
    }
    seen[i] = 1;
    for (int k = 0; k < numneigh[ii]; ++k) {
      const int j = firstneigh[ii][k];
      if (j < 0 || j >= nall) {
        set_err("neighbour %d of atom %d outside the %d atoms", j, i, nall);
        return false;
      }
      rows[(size_t)i * width + k] = j;
    }
  }
  for (int a = 0; a < nall; ++a)
    if (atype[a] < 0 || atype[a] >= dp->ntypes) {
      set_err("atom %d has type %d outside [0, %d)", a, atype[a], dp->ntypes);
      return false;
    }
  // slot order: atoms of the energy-carrying types grouped by type (stable)
  std::vector<int> order;
  dp->seg_type.clear();
  dp->seg_start.clear();
  dp->seg_end.clear();
  for (int t : dp->keep) {
    const int a0 = (int)order.size();
    for (int i = 0; i < nloc; ++i)
      if (atype[i] == t) order.push_back(i);
    dp->seg_type.push_back(t);
    dp->seg_start.push_back(a0);
    dp->seg_end.push_back((int)order.size());
  }
  dp->nreal = (int)order.size();
  const int S = dp->nnei, d3 = dp->d3, nr = dp->nreal > 0 ? dp->nreal : 1;
  const size_t nl1 = nloc > 0 ? nloc : 1;
  bool ok = dp->coord.reserve(sizeof(double) * 3 * nall) && dp->atype.reserve(sizeof(int) * nall) &&
            dp->nlist_in.reserve(sizeof(int) * rows.size()) &&
            dp->nlist.reserve(sizeof(int) * nl1 * S) && dp->status.reserve(sizeof(int) * 4) &&
            dp->em.reserve(sizeof(float) * nl1 * S * 4) && dp->xyz.reserve(sizeof(float) * nl1 * 4 * d3) &&
            dp->gxyz.reserve(sizeof(float) * nl1 * 4 * d3) && dp->atom_e.reserve(sizeof(double) * nl1) &&
            dp->nd.reserve(sizeof(float) * nl1 * S * 4) && dp->force.reserve(sizeof(double) * 3 * nall) &&
            dp->vrow.reserve(dpk_force_virial_workspace_bytes(nloc)) &&
            dp->gather_ws.reserve(dpk_force_virial_gather_workspace_bytes(nloc, S, nall)) &&
            dp->virial.reserve(sizeof(double) * 9) && dp->energy.reserve(sizeof(double)) &&
            dp->fsum_ws.reserve(dpk_frame_sum_workspace_bytes(1, nloc, 9)) &&
            dp->frame_of.reserve(sizeof(int) * nl1) && dp->frames.reserve(sizeof(dpk_frame)) &&
            dp->sums.reserve(sizeof(dpk_frame)) && dp->shift.reserve(sizeof(double) * 3) &&
            dp->order.reserve(sizeof(int) * nr) && dp->z1.reserve(sizeof(float) * nr * dp->h1) &&
            dp->gz.reserve(sizeof(unsigned short) * nr * 2 * dp->h1);
  if (!ok) return false;
  const dpk_frame fa = one_frame(nall), fl = one_frame(nloc);
  hipStream_t st = dp->stream;
  ok = hip_ok(hipMemcpyAsync(dp->nlist_in.p, rows.data(), sizeof(int) * rows.size(),
                             hipMemcpyHostToDevice, st), "list upload") &&
       hip_ok(hipMemcpyAsync(dp->atype.p, atype, sizeof(int) * nall, hipMemcpyHostToDevice, st),
              "type upload") &&
       hip_ok(hipMemcpyAsync(dp->frames.p, &fa, sizeof(fa), hipMemcpyHostToDevice, st), "frame") &&
       hip_ok(hipMemcpyAsync(dp->sums.p, &fl, sizeof(fl), hipMemcpyHostToDevice, st), "frame") &&
       hip_ok(hipMemsetAsync(dp->frame_of.p, 0, sizeof(int) * nl1, st), "memset") &&
       hip_ok(hipMemsetAsync(dp->shift.p, 0, sizeof(double) * 3, st), "memset") &&
       hip_ok(hipMemsetAsync(dp->status.p, 0, sizeof(int) * 4, st), "memset") &&
       (order.empty() || hip_ok(hipMemcpyAsync(dp->order.p, order.data(), sizeof(int) * order.size(),
                                                hipMemcpyHostToDevice, st), "order upload")) &&
       hip_ok(hipStreamSynchronize(st), "setup");  // host vectors go out of scope
  if (!ok) return false;
  dp->nloc = nloc;
  dp->nall = nall;
  dp->nnei_in = width;
  dp->have_list = true;
  return true;
}

bool step(dpk_deeppot* dp, const double* coord, bool check_list, double* energy, double* force,
          double* virial) {
  hipStream_t st = dp->stream;
  dpk_stream_t s = (dpk_stream_t)st;
  const int nloc = dp->nloc, nall = dp->nall, S = dp->nnei, d3 = dp->d3;
  if (!hip_ok(hipMemcpyAsync(dp->coord.p, coord, sizeof(double) * 3 * nall, hipMemcpyHostToDevice, st),
              "coordinate upload"))
    return false;
  const int* sel = dp->sel.data();
  const int* sec = dp->sec.data();
  const int* pm = static_cast<const int*>(bp(dp, "pair_mask"));
  const int* nl = dp->skip_padding ? dp->nlist.as<int>() : nullptr;
  const float* pad = nl ? static_cast<const float*>(bp(dp, "emb_pad")) : nullptr;
  if (!dpk_ok(dpk_nlist_format(dp->nlist.as<int>(), dp->status.as<int>(), dp->nlist_in.as<int>(),
                               dp->nnei_in, dp->coord.as<double>(), dp->atype.as<int>(), 1, nloc,
                               nall, dp->rcut, sel, dp->ntypes, 0, s),
              "dpk_nlist_format"))
    return false;
  if (nloc) {
    const dpk_frame* fr = dp->frames.as<dpk_frame>();
    if (!dpk_ok(dpk_env_mat_fused(dp->em.p, dp->prec, dp->coord.as<double>(), dp->atype.as<int>(),
                                  dp->frame_of.as<int>(), fr, dp->nlist.as<int>(),
                                  dp->shift.as<double>(), static_cast<const double*>(bp(dp, "davg")),
                                  static_cast<const double*>(bp(dp, "dstd")), nloc, S, dp->rcut,
                                  dp->rcut_smth, dp->protection, pm, dp->ntypes, s),
                "dpk_env_mat_fused") ||
        !dpk_ok(dpk_embed_fwd(dp->xyz.p, dp->prec, dp->em.p, dp->atype.as<int>(),
                              bp(dp, "emb_params"), bp(dp, "emb_image"),
                              static_cast<const int*>(bp(dp, "net_of")),
                              static_cast<const int*>(bp(dp, "ch_out")), dp->ntypes, dp->nnets,
                              dp->P1, dp->P2, dp->P3, dp->res2, dp->res3, dp->has_idt,
                              DPK_ACT_TANH, 0.0, d3, sec, nloc, S, nl, pad, s),
                "dpk_embed_fwd"))
      return false;
    if (dp->nreal < nloc &&
        !(hip_ok(hipMemsetAsync(dp->atom_e.p, 0, sizeof(double) * nloc, st), "memset") &&
          hip_ok(hipMemsetAsync(dp->gxyz.p, 0, sizeof(float) * nloc * 4 * d3, st), "memset")))
      return false;
    for (size_t k = 0; k < dp->seg_type.size(); ++k) {
      const int t = dp->seg_type[k], a = dp->seg_start[k], nt = dp->seg_end[k] - a;
      if (nt <= 0) continue;
      const int* ord = dp->order.as<int>() + a;
      float* z = dp->z1.as<float>() + (size_t)a * dp->h1;
      unsigned short* g = dp->gz.as<unsigned short>() + (size_t)a * 2 * dp->h1;
      const void* l1 = bp(dp, "l1_" + std::to_string(t));
      if (!dpk_ok(dpk_fit_l1_fwd(z, dp->fit_prec, dp->xyz.as<float>(), ord, nt, d3, dp->m2, l1,
                                 dp->h1, s),
                  "dpk_fit_l1_fwd") ||
          !dpk_ok(dpk_fit_tail_f16x2(dp->atom_e.as<double>(), g, z, bp(dp, "tail_" + std::to_string(t)),
                                     dp->layers, dp->tail_res[t], dp->ebias[t], ord, nt, s),
                  "dpk_fit_tail_f16x2") ||
          !dpk_ok(dpk_fit_l1_bwd(dp->gxyz.as<float>(), dp->fit_prec, g, dp->xyz.as<float>(), ord, nt,
                                 d3, dp->m2, l1, dp->h1, s),
                  "dpk_fit_l1_bwd"))
        return false;
    }
    if (!dpk_ok(dpk_embed_bwd(dp->nd.p, dp->prec, dp->em.p, dp->gxyz.p, dp->atype.as<int>(),
                              bp(dp, "emb_params"), bp(dp, "emb_image"),
                              static_cast<const int*>(bp(dp, "net_of")),
                              static_cast<const int*>(bp(dp, "ch_out")), dp->ntypes, dp->nnets,
                              dp->P1, dp->P2, dp->P3, dp->res2, dp->res3, dp->has_idt,
                              DPK_ACT_TANH, 0.0, d3, sec, nloc, S, nl, s),
                "dpk_embed_bwd"))
      return false;
  }
  if (!nloc &&
      !hip_ok(hipMemsetAsync(dp->force.p, 0, sizeof(double) * 3 * nall, st), "memset"))
    return false;
  if (nloc) {
    // neighbour forces gathered per extended atom (writes every force row)
    if (!dpk_ok(dpk_force_virial_gather(dp->force.as<double>(), nullptr,
                                        static_cast<double*>(dp->vrow.p), dp->nd.p, dp->prec,
                                        dp->coord.as<double>(), dp->atype.as<int>(),
                                        dp->nlist.as<int>(),
                                        static_cast<const double*>(bp(dp, "dstd")), nloc, S, nall,
                                        dp->rcut, dp->rcut_smth, dp->protection, pm, dp->ntypes,
                                        dp->gather_ws.p, dp->gather_ws.n, s),
                "dpk_force_virial_gather") ||
        !dpk_ok(dpk_frame_sum(dp->virial.as<double>(), static_cast<const double*>(dp->vrow.p), 9,
                              dp->sums.as<dpk_frame>(), 1, nloc, dp->fsum_ws.p, dp->fsum_ws.n, s),
                "dpk_frame_sum") ||
        !dpk_ok(dpk_frame_sum(dp->energy.as<double>(), dp->atom_e.as<double>(), 1,
                              dp->sums.as<dpk_frame>(), 1, nloc, dp->fsum_ws.p, dp->fsum_ws.n, s),
                "dpk_frame_sum"))
      return false;
  } else if (!(hip_ok(hipMemsetAsync(dp->virial.p, 0, sizeof(double) * 9, st), "memset") &&
               hip_ok(hipMemsetAsync(dp->energy.p, 0, sizeof(double), st), "memset"))) {
    return false;
  }
  int status[4] = {0, 0, 0, 0};
  if (!hip_ok(hipMemcpyAsync(energy, dp->energy.p, sizeof(double), hipMemcpyDeviceToHost, st), "copy") ||
      !hip_ok(hipMemcpyAsync(force, dp->force.p, sizeof(double) * 3 * nall, hipMemcpyDeviceToHost, st),
              "copy") ||
      !hip_ok(hipMemcpyAsync(virial, dp->virial.p, sizeof(double) * 9, hipMemcpyDeviceToHost, st), "copy") ||
      (check_list &&
       !hip_ok(hipMemcpyAsync(status, dp->status.p, sizeof(status), hipMemcpyDeviceToHost, st), "copy")) ||
      !hip_ok(hipStreamSynchronize(st), "step"))
    return false;
  if (check_list && status[1] != 0) {
    set_err("the neighbour list holds an index outside the extended region");
    return false;
  }
  return true;
}

}  // namespace

extern "C" dpk_deeppot* dpk_deeppot_load(const char* path, int device) {
  g_err[0] = 0;
  if (!hip_ok(hipSetDevice(device), "hipSetDevice")) return nullptr;
  dpk_deeppot* dp = new (std::nothrow) dpk_deeppot;
  if (!dp) {
    set_err("out of host memory");
    return nullptr;
  }
  dp->device = device;
  if (!hip_ok(hipStreamCreate(&dp->stream), "hipStreamCreate") || !load_file(dp, path) ||
      !check_model(dp)) {
    dpk_deeppot_free(dp);
    return nullptr;
  }
  return dp;
}

extern "C" void dpk_deeppot_free(dpk_deeppot* dp) {
  if (!dp) return;
  (void)hipSetDevice(dp->device);
  if (dp->stream) (void)hipStreamSynchronize(dp->stream);
  hipStream_t st = dp->stream;
  delete dp;  // device buffers first
  if (st) (void)hipStreamDestroy(st);
}

extern "C" int dpk_deeppot_compute(dpk_deeppot* dp, double* energy, double* force, double* virial,
                                   const double* coord, const int* atype, int nall, int nghost,
                                   int inum, const int* ilist, const int* numneigh,
                                   const int* const* firstneigh, int ago) {
  g_err[0] = 0;
  if (!dp || !energy || !force || !virial || !coord || !atype || nall < 0 || nghost < 0 ||
      nghost > nall) {
    set_err("dpk_deeppot_compute: bad arguments");
    return DPK_ERR_ARG;
  }
  if (!hip_ok(hipSetDevice(dp->device), "hipSetDevice")) return DPK_ERR_HIP;
  const int nloc = nall - nghost;
  const bool fresh = ago == 0 || !dp->have_list;
  if (fresh) {
    if (!ilist || !numneigh || (inum && !firstneigh)) {
      set_err("dpk_deeppot_compute: ago == 0 needs the neighbour list");
      return DPK_ERR_ARG;
    }
    if (!setup(dp, atype, nall, nloc, inum, ilist, numneigh, firstneigh)) return DPK_ERR_ARG;
  } else if (nall != dp->nall || nloc != dp->nloc) {
    set_err("dpk_deeppot_compute: ago > 0 with %d / %d atoms, the list has %d / %d", nloc, nall,
            dp->nloc, dp->nall);
    return DPK_ERR_ARG;
  }
  return step(dp, coord, fresh, energy, force, virial) ? DPK_OK : DPK_ERR_HIP;
}

extern "C" double dpk_deeppot_rcut(const dpk_deeppot* dp) { return dp ? dp->rcut : 0.0; }
extern "C" int dpk_deeppot_ntypes(const dpk_deeppot* dp) { return dp ? dp->ntypes : 0; }
extern "C" const char* dpk_deeppot_last_error(void) { return g_err; }
