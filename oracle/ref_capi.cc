// Plain C entry points over the reference's own C++ CPU kernels.
//
// TEST INFRASTRUCTURE ONLY: compiled by oracle/Makefile together with the
// reference sources where they lie (/root/reference/source/lib/src), into
// oracle/_ref/libdeepmd_ref.so.  Used by tests/ to pin the oracle and as the
// "reference" CPU baseline; never linked into the product.
//
// Wrapped reference functions (source/lib/include):
//   copy_coord_cpu            coord.h:25-34
//   build_nlist_cpu           neighbor_list.h:167-175
//   format_nlist_i_cpu        fmt_nlist.h:87-93
//   env_mat_a_cpu             env_mat.h:9-18
//   prod_env_mat_a_cpu        prod_env_mat.h:11-27
//   prod_force_a_cpu          prod_force.h:20-27
//   prod_virial_a_cpu         prod_virial.h:7-15
//   prod_force_grad_a_cpu     prod_force_grad.h:7-13
//   prod_virial_grad_a_cpu    prod_virial_grad.h:7-13
//   tabulate_fusion_se_a{,_grad,_grad_grad}_cpu   tabulate.h:29-72
#include <vector>

#include "coord.h"
#include "env_mat.h"
#include "fmt_nlist.h"
#include "neighbor_list.h"
#include "prod_env_mat.h"
#include "prod_force.h"
#include "prod_force_grad.h"
#include "prod_virial.h"
#include "prod_virial_grad.h"
#include "region.h"
#include "tabulate.h"

extern "C" {

// Periodic images by the reference's cell method. Returns nall, or -1 when
// mem_nall is too small.
int ref_copy_coord(double* out_c, int* out_t, int* mapping, const double* in_c,
                   const int* in_t, int nloc, int mem_nall, double rcut,
                   const double* box) {
  deepmd::Region<double> region;
  deepmd::init_region_cpu(region, box);
  int nall = 0;
  int ret = deepmd::copy_coord_cpu(out_c, out_t, mapping, &nall, in_c, in_t,
                                   nloc, mem_nall, (float)rcut, region);
  return ret == 0 ? nall : -1;
}

// Brute-force neighbour list over extended coordinates; rows of width
// max_nnei are filled with neighbour indices, counts written to numneigh.
int ref_build_nlist(int* nlist, int* numneigh, int max_nnei, const double* c_cpy,
                    int nloc, int nall, double rcut) {
  std::vector<int> ilist(nloc);
  std::vector<int*> first(nloc);
  for (int ii = 0; ii < nloc; ++ii) {
    ilist[ii] = ii;
    first[ii] = nlist + (size_t)ii * max_nnei;
  }
  deepmd::InputNlist in(nloc, ilist.data(), numneigh, first.data());
  int max_size = 0;
  return deepmd::build_nlist_cpu(in, &max_size, c_cpy, nloc, nall, max_nnei,
                                 (float)rcut);
}

// Format one atom's raw neighbour list (type blocks, FP32 distance keys).
int ref_format_nlist_i(int* fmt, const double* posi, const int* type, int nall,
                       int i_idx, const int* nei, int n_nei, double rcut,
                       const int* sec, int nsec) {
  std::vector<int> out;
  std::vector<double> p(posi, posi + (size_t)nall * 3);
  std::vector<int> t(type, type + nall);
  std::vector<int> ni(nei, nei + n_nei);
  std::vector<int> s(sec, sec + nsec);
  int ret = format_nlist_i_cpu(out, p, t, i_idx, ni, (float)rcut, s);
  for (size_t k = 0; k < out.size(); ++k) fmt[k] = out[k];
  return ret;
}

// Un-normalised env matrix + derivative + rij for every local atom from a
// formatted nlist [nloc, nnei].
void ref_env_mat_a(double* em, double* em_deriv, double* rij,
                   const double* posi, const int* type, int nloc, int nall,
                   const int* fmt_nlist, const int* sec, int nsec,
                   double rcut_smth, double rcut) {
  std::vector<double> p(posi, posi + (size_t)nall * 3);
  std::vector<int> t(type, type + nall);
  std::vector<int> s(sec, sec + nsec);
  int nnei = sec[nsec - 1];
  for (int ii = 0; ii < nloc; ++ii) {
    std::vector<int> f(fmt_nlist + (size_t)ii * nnei,
                       fmt_nlist + (size_t)(ii + 1) * nnei);
    std::vector<double> e, d, r;
    deepmd::env_mat_a_cpu(e, d, r, p, t, ii, f, s, (float)rcut_smth,
                          (float)rcut);
    for (int k = 0; k < nnei * 4; ++k) em[(size_t)ii * nnei * 4 + k] = e[k];
    for (int k = 0; k < nnei * 12; ++k)
      em_deriv[(size_t)ii * nnei * 12 + k] = d[k];
    for (int k = 0; k < nnei * 3; ++k) rij[(size_t)ii * nnei * 3 + k] = r[k];
  }
}

// Full reference op: format raw lists, env matrix, normalisation.
void ref_prod_env_mat_a(double* em, double* em_deriv, double* rij, int* nlist,
                        const double* coord, const int* type,
                        const int* raw_nlist, const int* raw_numneigh,
                        int raw_width, const double* avg, const double* stdv,
                        int nloc, int nall, double rcut, double rcut_smth,
                        const int* sec, int nsec) {
  std::vector<int> ilist(nloc);
  std::vector<int> numneigh(raw_numneigh, raw_numneigh + nloc);
  std::vector<int*> first(nloc);
  for (int ii = 0; ii < nloc; ++ii) {
    ilist[ii] = ii;
    first[ii] = const_cast<int*>(raw_nlist) + (size_t)ii * raw_width;
  }
  deepmd::InputNlist in(nloc, ilist.data(), numneigh.data(), first.data());
  std::vector<int> s(sec, sec + nsec);
  deepmd::prod_env_mat_a_cpu(em, em_deriv, rij, nlist, coord, type, in,
                             raw_width, avg, stdv, nloc, nall, (float)rcut,
                             (float)rcut_smth, s);
}

void ref_prod_force_a(double* force, const double* net_deriv,
                      const double* in_deriv, const int* nlist, int nloc,
                      int nall, int nnei, int nframes) {
  deepmd::prod_force_a_cpu(force, net_deriv, in_deriv, nlist, nloc, nall, nnei,
                           nframes);
}

void ref_prod_virial_a(double* virial, double* atom_virial,
                       const double* net_deriv, const double* in_deriv,
                       const double* rij, const int* nlist, int nloc, int nall,
                       int nnei) {
  deepmd::prod_virial_a_cpu(virial, atom_virial, net_deriv, in_deriv, rij,
                            nlist, nloc, nall, nnei);
}

void ref_prod_force_grad_a(double* grad_net, const double* grad,
                           const double* env_deriv, const int* nlist, int nloc,
                           int nnei, int nframes) {
  deepmd::prod_force_grad_a_cpu(grad_net, grad, env_deriv, nlist, nloc, nnei,
                                nframes);
}

void ref_prod_virial_grad_a(double* grad_net, const double* grad,
                            const double* env_deriv, const double* rij,
                            const int* nlist, int nloc, int nnei) {
  deepmd::prod_virial_grad_a_cpu(grad_net, grad, env_deriv, rij, nlist, nloc,
                                 nnei);
}

#define DPK_REF_TAB(T, SFX)                                                    \
  void ref_tabulate_fusion_se_a_##SFX(T* out, const T* table, const T* info,   \
                                      const T* em_x, const T* em, int nloc,    \
                                      int nnei, int L, int ndescrpt) {         \
    deepmd::tabulate_fusion_se_a_cpu<T>(out, table, info, em_x, em, nullptr,   \
                                        nloc, nnei, L, true, ndescrpt);        \
  }                                                                            \
  void ref_tabulate_fusion_se_a_grad_##SFX(                                    \
      T* dy_dem_x, T* dy_dem, const T* table, const T* info, const T* em_x,    \
      const T* em, const T* dy, int nloc, int nnei, int L, int ndescrpt) {     \
    deepmd::tabulate_fusion_se_a_grad_cpu<T>(dy_dem_x, dy_dem, nullptr, table, \
                                             info, em_x, em, nullptr, dy,      \
                                             nloc, nnei, L, true, ndescrpt);   \
  }                                                                            \
  void ref_tabulate_fusion_se_a_grad_grad_##SFX(                               \
      T* dz_dy, const T* table, const T* info, const T* em_x, const T* em,     \
      const T* dz_dy_dem_x, const T* dz_dy_dem, int nloc, int nnei, int L,     \
      int ndescrpt) {                                                          \
    deepmd::tabulate_fusion_se_a_grad_grad_cpu<T>(                             \
        dz_dy, table, info, em_x, em, nullptr, dz_dy_dem_x, dz_dy_dem,         \
        nullptr, nloc, nnei, L, true, ndescrpt);                               \
  }

DPK_REF_TAB(double, f64)
DPK_REF_TAB(float, f32)

}  // extern "C"
