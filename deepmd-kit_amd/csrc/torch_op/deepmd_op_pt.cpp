// libdeepmd_op_pt.so -- the torch.ops.deepmd operator library of the se_e2_a
// path on MI355X, loadable with torch.ops.load_library exactly as the
// reference's op library (deepmd/pt/cxx_op.py:42-56) and usable from a C++
// libtorch process (TorchScript models calling torch.ops.deepmd.*).
//
// Registration mirrors source/op/pt: TORCH_LIBRARY(deepmd) defines the
// enable_mpi sentinel (print_summary.cc:5-14) and the ops are added with
// TORCH_LIBRARY_FRAGMENT (tabulate_multi_device.cc:1682-1684).  Every op is
// a host wrapper over the C ABI of libdpk (include/dpk.h): shape / dtype /
// device checks on the host, kernels enqueued on the current HIP stream, no
// device synchronisation and no reads of device data on the host, so the ops
// can be captured in a HIP graph.  Only HIP tensors are computed; a CPU
// tensor raises after the same validation the reference performs.
#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>
#include <torch/autograd.h>
#include <torch/library.h>
#include <torch/torch.h>

#include <stdexcept>
#include <string>
#include <utility>

#include "dpk.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void* stream_of(const Tensor& t) {
  return reinterpret_cast<void*>(c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

void check_rc(int rc, const char* what) {
  TORCH_CHECK(rc == DPK_OK, what, " failed (status ", rc, "): ", dpk_last_error());
}

void require_gpu(const Tensor& t, const char* op) {
  if (!t.is_cuda())
    throw std::runtime_error(std::string("deepmd::") + op +
                             ": the input tensor is on the CPU; deepmd_amd implements this "
                             "operator for MI355X (HIP) tensors only");
}

// ------------------------------------------------------------- tabulate --
// Checks of source/op/pt/tabulate_multi_device.cc:21-50 (device contract,
// basis dimension) and :72-83 (ranks), plus dtype and size consistency of
// every raw-pointer operand (each must match table's dtype and the shapes
// the kernel reads).

using Named = std::pair<const char*, Tensor>;

void check_tab_devices(const Tensor& table, const Tensor& info, std::initializer_list<Named> ts) {
  TORCH_CHECK(info.device().is_cpu(), "table_info must be on the CPU");
  for (const auto& nt : ts) {
    if (!nt.second.defined()) continue;
    TORCH_CHECK(nt.second.device() == table.device(), nt.first,
                " must be on the same device as table; table is on ", table.device(), " but ",
                nt.first, " is on ", nt.second.device());
  }
}

void check_tab_dtypes(const Tensor& table, std::initializer_list<Named> ts) {
  TORCH_CHECK(table.scalar_type() == torch::kFloat64 || table.scalar_type() == torch::kFloat32,
              "tabulate: unsupported table dtype ", table.scalar_type(), " (float or double)");
  for (const auto& nt : ts) {
    if (!nt.second.defined()) continue;
    TORCH_CHECK(nt.second.scalar_type() == table.scalar_type(), nt.first,
                " must have the dtype of table (", table.scalar_type(), "), got ",
                nt.second.scalar_type());
  }
}

void check_nd(int64_t nd) {
  TORCH_CHECK(nd == 4 || nd == 9 || nd == 16 || nd == 25,
              "The environment basis dimension must be 4, 9, 16, or 25, got ", nd);
}

struct TabArgs {
  Tensor table, info, em_x, em;
  int64_t nloc, nnei, nd, L;
  bool f64;
};

TabArgs tab_common(const Tensor& table, const Tensor& info, const Tensor& em_x, const Tensor& em,
                   int64_t L, std::initializer_list<Named> more, const char* op) {
  if (table.dim() != 2) throw std::invalid_argument("Dim of table should be 2");
  if (em_x.dim() != 2) throw std::invalid_argument("Dim of input should be 2");
  if (em.dim() != 3) throw std::invalid_argument("Dim of input should be 3");
  TabArgs a;
  a.nloc = em.size(0);
  a.nnei = em.size(1);
  a.nd = em.size(2);
  a.L = L;
  check_nd(a.nd);
  check_tab_dtypes(table, {{"em_x", em_x}, {"em", em}});
  check_tab_dtypes(table, more);
  TORCH_CHECK(em_x.numel() == a.nloc * a.nnei, op, ": em_x holds ", em_x.numel(),
              " values, em has ", a.nloc, " x ", a.nnei, " slots");
  TORCH_CHECK(L > 0 && table.size(1) >= 6 * L, op, ": table rows of ", table.size(1),
              " values cannot hold ", L, " channels x 6 coefficients");
  TORCH_CHECK(info.numel() >= 6, op, ": table_info needs 6 values");
  a.f64 = table.scalar_type() == torch::kFloat64;
  a.table = table.contiguous();
  a.info = info.to(table.scalar_type()).contiguous();
  a.em_x = em_x.contiguous();
  a.em = em.contiguous();
  return a;
}

Tensor tab_fwd(const Tensor& table, const Tensor& info, const Tensor& em_x, const Tensor& em,
               int64_t L) {
  check_tab_devices(table, info, {{"em_x", em_x}, {"em", em}});
  TabArgs a = tab_common(table, info, em_x, em, L, {}, "tabulate_fusion_se_a");
  require_gpu(table, "tabulate_fusion_se_a");
  const c10::DeviceGuard guard(table.device());
  Tensor out = torch::empty({a.nloc, a.nd, L}, table.options());
  if (a.nloc == 0) return out;
  void* st = stream_of(table);
  int rc = a.f64 ? dpk_tabulate_fusion_se_a_f64(
                       out.data_ptr<double>(), a.table.data_ptr<double>(), (int)table.size(0),
                       a.info.data_ptr<double>(), a.em_x.data_ptr<double>(),
                       a.em.data_ptr<double>(), (int)a.nloc, (int)a.nnei, (int)L, (int)a.nd, st)
                 : dpk_tabulate_fusion_se_a_f32(
                       out.data_ptr<float>(), a.table.data_ptr<float>(), (int)table.size(0),
                       a.info.data_ptr<float>(), a.em_x.data_ptr<float>(),
                       a.em.data_ptr<float>(), (int)a.nloc, (int)a.nnei, (int)L, (int)a.nd, st);
  check_rc(rc, "tabulate_fusion_se_a");
  return out;
}

std::vector<Tensor> tab_grad(const Tensor& table, const Tensor& info, const Tensor& em_x,
                             const Tensor& em, const Tensor& dy, const Tensor& descriptor) {
  check_tab_devices(table, info,
                    {{"em_x", em_x}, {"em", em}, {"dy", dy}, {"descriptor", descriptor}});
  if (dy.dim() != 3) throw std::invalid_argument("Dim of dy_tensor should be 3");
  TORCH_CHECK(descriptor.dim() == 3, "tabulate_fusion_se_a_grad: descriptor must be 3-d");
  const int64_t L = descriptor.size(2);
  TabArgs a = tab_common(table, info, em_x, em, L, {{"dy", dy}}, "tabulate_fusion_se_a_grad");
  TORCH_CHECK(dy.numel() == a.nloc * a.nd * L, "tabulate_fusion_se_a_grad: dy holds ",
              dy.numel(), " values, expected ", a.nloc, " x ", a.nd, " x ", L);
  require_gpu(table, "tabulate_fusion_se_a_grad");
  const c10::DeviceGuard guard(table.device());
  Tensor dx = torch::zeros_like(a.em_x);
  Tensor dem = torch::zeros_like(a.em);
  if (a.nloc == 0) return {dx, dem};
  Tensor dyc = dy.contiguous();
  void* st = stream_of(table);
  int rc = a.f64 ? dpk_tabulate_fusion_se_a_grad_f64(
                       dx.data_ptr<double>(), dem.data_ptr<double>(), a.table.data_ptr<double>(),
                       (int)table.size(0), a.info.data_ptr<double>(), a.em_x.data_ptr<double>(),
                       a.em.data_ptr<double>(), dyc.data_ptr<double>(), (int)a.nloc,
                       (int)a.nnei, (int)L, (int)a.nd, st)
                 : dpk_tabulate_fusion_se_a_grad_f32(
                       dx.data_ptr<float>(), dem.data_ptr<float>(), a.table.data_ptr<float>(),
                       (int)table.size(0), a.info.data_ptr<float>(), a.em_x.data_ptr<float>(),
                       a.em.data_ptr<float>(), dyc.data_ptr<float>(), (int)a.nloc, (int)a.nnei,
                       (int)L, (int)a.nd, st);
  check_rc(rc, "tabulate_fusion_se_a_grad");
  return {dx.view(em_x.sizes()), dem.view(em.sizes())};
}

Tensor tab_grad_grad(const Tensor& table, const Tensor& info, const Tensor& em_x,
                     const Tensor& em, const Tensor& dzx, const Tensor& dzem,
                     const Tensor& descriptor) {
  check_tab_devices(table, info,
                    {{"em_x", em_x},
                     {"em", em},
                     {"dz_dy_dem_x", dzx},
                     {"dz_dy_dem", dzem},
                     {"descriptor", descriptor}});
  if (dzx.dim() != 2) throw std::invalid_argument("Dim of dz_dy_dem_x should be 2");
  if (dzem.dim() != 3) throw std::invalid_argument("Dim of dz_dy_dem should be 3");
  TORCH_CHECK(descriptor.dim() == 3, "tabulate_fusion_se_a_grad_grad: descriptor must be 3-d");
  const int64_t L = descriptor.size(2);
  TabArgs a = tab_common(table, info, em_x, em, L,
                         {{"dz_dy_dem_x", dzx}, {"dz_dy_dem", dzem}},
                         "tabulate_fusion_se_a_grad_grad");
  TORCH_CHECK(L <= 1024,
              "In the process of model compression, the size of the last layer of embedding "
              "net must be less than 1024!");
  TORCH_CHECK(dzx.numel() == a.nloc * a.nnei && dzem.numel() == a.nloc * a.nnei * a.nd,
              "tabulate_fusion_se_a_grad_grad: dz_dy_dem_x / dz_dy_dem do not match em");
  require_gpu(table, "tabulate_fusion_se_a_grad_grad");
  const c10::DeviceGuard guard(table.device());
  Tensor out = torch::empty({a.nloc, a.nd, L}, table.options());
  if (a.nloc == 0) return out;
  Tensor zx = dzx.contiguous(), ze = dzem.contiguous();
  void* st = stream_of(table);
  int rc = a.f64 ? dpk_tabulate_fusion_se_a_grad_grad_f64(
                       out.data_ptr<double>(), a.table.data_ptr<double>(), (int)table.size(0),
                       a.info.data_ptr<double>(), a.em_x.data_ptr<double>(),
                       a.em.data_ptr<double>(), zx.data_ptr<double>(), ze.data_ptr<double>(),
                       (int)a.nloc, (int)a.nnei, (int)L, (int)a.nd, st)
                 : dpk_tabulate_fusion_se_a_grad_grad_f32(
                       out.data_ptr<float>(), a.table.data_ptr<float>(), (int)table.size(0),
                       a.info.data_ptr<float>(), a.em_x.data_ptr<float>(),
                       a.em.data_ptr<float>(), zx.data_ptr<float>(), ze.data_ptr<float>(),
                       (int)a.nloc, (int)a.nnei, (int)L, (int)a.nd, st);
  check_rc(rc, "tabulate_fusion_se_a_grad_grad");
  return out;
}

// Autograd chain of tabulate_multi_device.cc:731-962: forward -> grad op,
// whose backward is grad_grad (gradient with respect to dy only).
class TabulateFusionSeAGradOp : public torch::autograd::Function<TabulateFusionSeAGradOp> {
 public:
  static variable_list forward(AutogradContext* ctx, const Tensor& table, const Tensor& info,
                               const Tensor& em_x, const Tensor& em, const Tensor& dy,
                               const Tensor& descriptor) {
    auto r = tab_grad(table, info, em_x, em, dy, descriptor);
    ctx->save_for_backward({table, info, em_x, em, descriptor});
    return {r[0], r[1]};
  }
  static variable_list backward(AutogradContext* ctx, variable_list g) {
    auto s = ctx->get_saved_variables();
    Tensor gx = g[0].defined() ? g[0] : torch::zeros_like(s[2]);
    Tensor gm = g[1].defined() ? g[1] : torch::zeros_like(s[3]);
    Tensor dz = tab_grad_grad(s[0], s[1], s[2], s[3], gx, gm, s[4]);
    return {Tensor(), Tensor(), Tensor(), Tensor(), dz, Tensor()};
  }
};

class TabulateFusionSeAOp : public torch::autograd::Function<TabulateFusionSeAOp> {
 public:
  static variable_list forward(AutogradContext* ctx, const Tensor& table, const Tensor& info,
                               const Tensor& em_x, const Tensor& em, int64_t L) {
    Tensor out = tab_fwd(table, info, em_x, em, L);
    ctx->save_for_backward({table, info, em_x, em, out});
    return {out};
  }
  static variable_list backward(AutogradContext* ctx, variable_list g) {
    auto s = ctx->get_saved_variables();
    auto r = TabulateFusionSeAGradOp::apply(s[0], s[1], s[2], s[3], g[0].contiguous(), s[4]);
    return {Tensor(), Tensor(), r[0], r[1], Tensor()};
  }
};

std::vector<Tensor> tabulate_fusion_se_a(const Tensor& table, const Tensor& table_info,
                                         const Tensor& em_x, const Tensor& em,
                                         int64_t last_layer_size) {
  return TabulateFusionSeAOp::apply(table, table_info, em_x, em, last_layer_size);
}

std::vector<Tensor> tabulate_fusion_se_a_grad(const Tensor& table, const Tensor& table_info,
                                              const Tensor& em_x, const Tensor& em,
                                              const Tensor& dy, const Tensor& descriptor) {
  return TabulateFusionSeAGradOp::apply(table, table_info, em_x, em, dy, descriptor);
}

Tensor tabulate_fusion_se_a_grad_grad(const Tensor& table, const Tensor& table_info,
                                      const Tensor& em_x, const Tensor& em,
                                      const Tensor& dz_dy_dem_x, const Tensor& dz_dy_dem,
                                      const Tensor& descriptor) {
  return tab_grad_grad(table, table_info, em_x, em, dz_dy_dem_x, dz_dy_dem, descriptor);
}

// ------------------------------------------------ env mat / force / virial --
// Torch forms of the TensorFlow ops (source/op/tf/prod_env_mat_multi_device.cc,
// prod_force_multi_device.cc, prod_virial_multi_device.cc and their grads)
// computed in FP64 (FP32 for an FP32 env matrix); results take the dtype of
// the first input.

Tensor f64c(const Tensor& t) { return t.to(torch::kFloat64).contiguous(); }
Tensor i32c(const Tensor& t) { return t.to(torch::kInt32).contiguous(); }

void same_device(const char* op, std::initializer_list<Named> ts) {
  const Tensor* first = nullptr;
  for (const auto& nt : ts) {
    require_gpu(nt.second, op);
    if (first == nullptr) first = &nt.second;
    TORCH_CHECK(nt.second.device() == first->device(), op, ": ", nt.first,
                " is on another device");
  }
}

int64_t slots_per_row(const Tensor& nlist, int64_t nf, int64_t nloc, const char* op) {
  if (nf * nloc == 0) return 0;
  TORCH_CHECK(nlist.numel() % (nf * nloc) == 0, op, ": nlist of ", nlist.numel(),
              " entries is not [nframes, nloc, nnei] with nloc = ", nloc);
  return nlist.numel() / (nf * nloc);
}

Tensor force_impl(const Tensor& net_deriv, const Tensor& in_deriv, const Tensor& nlist,
                  int64_t nloc, int64_t nall) {
  const char* op = "prod_force_se_a";
  same_device(op, {{"net_deriv", net_deriv}, {"in_deriv", in_deriv}, {"nlist", nlist}});
  TORCH_CHECK(nloc >= 0 && nall >= nloc, op, ": nall must be >= nloc");
  const int64_t nf = net_deriv.size(0);
  const int64_t nnei = slots_per_row(nlist, nf, nloc, op);
  TORCH_CHECK(net_deriv.numel() == nf * nloc * nnei * 4 && in_deriv.numel() == nf * nloc * nnei * 12,
              op, ": net_deriv / in_deriv do not match nlist");
  const c10::DeviceGuard guard(net_deriv.device());
  Tensor force = torch::zeros({nf, nall * 3}, net_deriv.options().dtype(torch::kFloat64));
  if (nf * nloc * nnei == 0) return force.to(net_deriv.scalar_type());
  Tensor g = f64c(net_deriv), d = f64c(in_deriv), n = i32c(nlist);
  check_rc(dpk_prod_force_a_f64(force.data_ptr<double>(), g.data_ptr<double>(),
                                d.data_ptr<double>(), n.data_ptr<int>(), (int)nf, (int)nloc,
                                (int)nall, (int)nnei, stream_of(net_deriv)),
           op);
  return force.to(net_deriv.scalar_type());
}

Tensor force_grad_impl(const Tensor& grad, const Tensor& in_deriv, const Tensor& nlist,
                       int64_t nloc, int64_t ngrad, c10::ScalarType out_dtype) {
  const char* op = "prod_force_se_a_grad";
  same_device(op, {{"grad", grad}, {"in_deriv", in_deriv}, {"nlist", nlist}});
  const int64_t nf = grad.size(0);
  const int64_t nnei = slots_per_row(nlist, nf, nloc, op);
  TORCH_CHECK(grad.numel() == nf * ngrad * 3 && in_deriv.numel() == nf * nloc * nnei * 12, op,
              ": grad / in_deriv do not match nlist");
  const c10::DeviceGuard guard(grad.device());
  Tensor gn = torch::zeros({nf, nloc * nnei * 4}, grad.options().dtype(torch::kFloat64));
  if (nf * nloc * nnei == 0) return gn.to(out_dtype);
  Tensor g = f64c(grad), d = f64c(in_deriv), n = i32c(nlist);
  check_rc(dpk_prod_force_grad_a_f64(gn.data_ptr<double>(), g.data_ptr<double>(),
                                     d.data_ptr<double>(), n.data_ptr<int>(), (int)nf,
                                     (int)nloc, (int)nnei, (int)ngrad, stream_of(grad)),
           op);
  return gn.to(out_dtype);
}

Tensor virial_grad_impl(const Tensor& grad, const Tensor& in_deriv, const Tensor& rij,
                        const Tensor& nlist, int64_t nloc, c10::ScalarType out_dtype) {
  const char* op = "prod_virial_se_a_grad";
  same_device(op, {{"grad", grad}, {"in_deriv", in_deriv}, {"rij", rij}, {"nlist", nlist}});
  const int64_t nf = grad.size(0);
  const int64_t nnei = slots_per_row(nlist, nf, nloc, op);
  TORCH_CHECK(grad.numel() == nf * 9 && in_deriv.numel() == nf * nloc * nnei * 12 &&
                  rij.numel() == nf * nloc * nnei * 3,
              op, ": grad / in_deriv / rij do not match nlist");
  const c10::DeviceGuard guard(grad.device());
  Tensor gn = torch::zeros({nf, nloc * nnei * 4}, grad.options().dtype(torch::kFloat64));
  if (nf * nloc * nnei == 0) return gn.to(out_dtype);
  Tensor g = f64c(grad), d = f64c(in_deriv), r = f64c(rij), n = i32c(nlist);
  check_rc(dpk_prod_virial_grad_a_f64(gn.data_ptr<double>(), g.data_ptr<double>(),
                                      d.data_ptr<double>(), r.data_ptr<double>(),
                                      n.data_ptr<int>(), (int)nf, (int)nloc, (int)nnei,
                                      stream_of(grad)),
           op);
  return gn.to(out_dtype);
}

std::vector<Tensor> virial_impl(const Tensor& net_deriv, const Tensor& in_deriv, const Tensor& rij,
                                const Tensor& nlist, int64_t nloc, int64_t nall) {
  const char* op = "prod_virial_se_a";
  same_device(op, {{"net_deriv", net_deriv}, {"in_deriv", in_deriv}, {"rij", rij},
                   {"nlist", nlist}});
  TORCH_CHECK(nloc >= 0 && nall >= nloc, op, ": nall must be >= nloc");
  const int64_t nf = net_deriv.size(0);
  const int64_t nnei = slots_per_row(nlist, nf, nloc, op);
  TORCH_CHECK(net_deriv.numel() == nf * nloc * nnei * 4 &&
                  in_deriv.numel() == nf * nloc * nnei * 12 && rij.numel() == nf * nloc * nnei * 3,
              op, ": net_deriv / in_deriv / rij do not match nlist");
  const c10::DeviceGuard guard(net_deriv.device());
  auto o64 = net_deriv.options().dtype(torch::kFloat64);
  Tensor vir = torch::zeros({nf, 9}, o64);
  Tensor av = torch::zeros({nf, nall * 9}, o64);
  if (nf * nloc * nnei > 0) {
    Tensor ws = torch::empty({(int64_t)dpk_prod_virial_workspace_bytes((int)nf, (int)nloc) + 8},
                             net_deriv.options().dtype(torch::kUInt8));
    Tensor g = f64c(net_deriv), d = f64c(in_deriv), r = f64c(rij), n = i32c(nlist);
    check_rc(dpk_prod_virial_a_f64(vir.data_ptr<double>(), av.data_ptr<double>(),
                                   g.data_ptr<double>(), d.data_ptr<double>(),
                                   r.data_ptr<double>(), n.data_ptr<int>(), (int)nf, (int)nloc,
                                   (int)nall, (int)nnei, ws.data_ptr(), (size_t)ws.numel(),
                                   stream_of(net_deriv)),
             op);
  }
  return {vir.to(net_deriv.scalar_type()), av.to(net_deriv.scalar_type())};
}

std::vector<Tensor> env_impl(const Tensor& coord, const Tensor& atype, const Tensor& nlist,
                             const Tensor& davg, const Tensor& dstd, double rcut,
                             double rcut_smth, double protection) {
  const char* op = "prod_env_mat_a";
  same_device(op, {{"coord", coord}, {"atype", atype}, {"nlist", nlist}, {"davg", davg},
                   {"dstd", dstd}});
  TORCH_CHECK(nlist.dim() == 3, op, ": nlist must be [nframes, nloc, nnei]");
  TORCH_CHECK(atype.dim() == 2, op, ": atype must be [nframes, nall]");
  const int64_t nf = nlist.size(0), nloc = nlist.size(1), nnei = nlist.size(2);
  const int64_t nall = atype.size(1);
  TORCH_CHECK(atype.size(0) == nf && nall >= nloc, op, ": atype does not match nlist");
  TORCH_CHECK(coord.numel() == nf * nall * 3, op, ": coord must hold nframes x nall x 3 values");
  TORCH_CHECK(nnei > 0 && davg.numel() > 0 && davg.numel() % (nnei * 4) == 0 &&
                  dstd.numel() == davg.numel(),
              op, ": davg / dstd must be [ntypes, nnei * 4]");
  const auto dt = coord.scalar_type();
  TORCH_CHECK(dt == torch::kFloat64 || dt == torch::kFloat32, op, ": unsupported dtype ", dt);
  const int ntypes = (int)(davg.numel() / (nnei * 4));
  const c10::DeviceGuard guard(coord.device());
  auto opt = coord.options();
  Tensor em = torch::empty({nf, nloc, nnei * 4}, opt);
  Tensor ed = torch::empty({nf, nloc, nnei * 12}, opt);
  Tensor rij = torch::empty({nf, nloc, nnei * 3}, opt);
  if (nf * nloc == 0) return {em, ed, rij};
  Tensor c = coord.reshape({nf, nall * 3}).contiguous(), t = i32c(atype), n = i32c(nlist);
  Tensor a = davg.to(dt).contiguous(), s = dstd.to(dt).contiguous();
  void* st = stream_of(coord);
  int rc = dt == torch::kFloat64
               ? dpk_prod_env_mat_a_f64(em.data_ptr<double>(), ed.data_ptr<double>(),
                                        rij.data_ptr<double>(), c.data_ptr<double>(),
                                        t.data_ptr<int>(), n.data_ptr<int>(), a.data_ptr<double>(),
                                        s.data_ptr<double>(), (int)nf, (int)nloc, (int)nall,
                                        (int)nnei, ntypes, rcut, rcut_smth, protection, st)
               : dpk_prod_env_mat_a_f32(em.data_ptr<float>(), ed.data_ptr<float>(),
                                        rij.data_ptr<float>(), c.data_ptr<float>(),
                                        t.data_ptr<int>(), n.data_ptr<int>(), a.data_ptr<float>(),
                                        s.data_ptr<float>(), (int)nf, (int)nloc, (int)nall,
                                        (int)nnei, ntypes, rcut, rcut_smth, protection, st);
  check_rc(rc, op);
  return {em, ed, rij};
}

// d E / d coord from the gradients of em (through em_deriv: the centre gets
// -g.dR, the neighbour +g.dR, i.e. minus prod_force) and of rij = r_j - r_i.
// em_deriv is not differentiable (its gradient is ignored).
class ProdEnvMatAOp : public torch::autograd::Function<ProdEnvMatAOp> {
 public:
  static variable_list forward(AutogradContext* ctx, const Tensor& coord, const Tensor& atype,
                               const Tensor& nlist, const Tensor& davg, const Tensor& dstd,
                               double rcut, double rcut_smth, double protection) {
    auto r = env_impl(coord, atype, nlist, davg, dstd, rcut, rcut_smth, protection);
    ctx->set_materialize_grads(false);
    ctx->mark_non_differentiable({r[1]});
    ctx->save_for_backward({r[1], i32c(nlist)});
    ctx->saved_data["nall"] = atype.size(1);
    ctx->saved_data["coord_sizes"] = coord.sizes().vec();
    return r;
  }
  static variable_list backward(AutogradContext* ctx, variable_list g) {
    auto s = ctx->get_saved_variables();
    const Tensor& ed = s[0];
    const Tensor& nl = s[1];
    const int64_t nall = ctx->saved_data["nall"].toInt();
    auto csz = ctx->saved_data["coord_sizes"].toIntVector();
    const int64_t nf = nl.size(0), nloc = nl.size(1);
    Tensor gc;
    if (g[0].defined()) gc = -force_impl(g[0].contiguous(), ed, nl, nloc, nall);
    if (g[2].defined()) {
      Tensor gr = g[2].reshape({nf, nloc, -1, 3}).to(torch::kFloat64);
      Tensor m = (nl >= 0).unsqueeze(-1);
      Tensor gm = torch::where(m, gr, torch::zeros_like(gr));
      Tensor acc = torch::zeros({nf, nall, 3}, gr.options());
      acc.narrow(1, 0, nloc).sub_(gm.sum(2));
      Tensor idx = torch::where(nl >= 0, nl, torch::zeros_like(nl)).to(torch::kInt64);
      acc.scatter_add_(1, idx.reshape({nf, -1, 1}).expand({nf, idx.size(1) * idx.size(2), 3}),
                       gm.reshape({nf, -1, 3}));
      Tensor add = acc.reshape({nf, nall * 3});
      gc = gc.defined() ? gc + add.to(gc.scalar_type()) : add;
    }
    if (gc.defined()) gc = gc.reshape(csz).to(ed.scalar_type());
    return {gc, Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
  }
};

class ProdForceSeAOp : public torch::autograd::Function<ProdForceSeAOp> {
 public:
  static variable_list forward(AutogradContext* ctx, const Tensor& net_deriv,
                               const Tensor& in_deriv, const Tensor& nlist, int64_t nloc,
                               int64_t nall) {
    ctx->save_for_backward({in_deriv, nlist});
    ctx->saved_data["nloc"] = nloc;
    ctx->saved_data["nall"] = nall;
    ctx->saved_data["nd_sizes"] = net_deriv.sizes().vec();
    ctx->saved_data["nd_dtype"] = (int64_t)net_deriv.scalar_type();
    return {force_impl(net_deriv, in_deriv, nlist, nloc, nall)};
  }
  static variable_list backward(AutogradContext* ctx, variable_list g) {
    auto s = ctx->get_saved_variables();
    Tensor gn = force_grad_impl(g[0].contiguous(), s[0], s[1], ctx->saved_data["nloc"].toInt(),
                                ctx->saved_data["nall"].toInt(),
                                (c10::ScalarType)ctx->saved_data["nd_dtype"].toInt());
    return {gn.view(ctx->saved_data["nd_sizes"].toIntVector()), Tensor(), Tensor(), Tensor(),
            Tensor()};
  }
};

class ProdVirialSeAOp : public torch::autograd::Function<ProdVirialSeAOp> {
 public:
  static variable_list forward(AutogradContext* ctx, const Tensor& net_deriv,
                               const Tensor& in_deriv, const Tensor& rij, const Tensor& nlist,
                               int64_t nloc, int64_t nall) {
    auto r = virial_impl(net_deriv, in_deriv, rij, nlist, nloc, nall);
    ctx->mark_non_differentiable({r[1]});
    ctx->save_for_backward({net_deriv, in_deriv, rij, nlist});
    ctx->saved_data["nloc"] = nloc;
    return r;
  }
  static variable_list backward(AutogradContext* ctx, variable_list g) {
    auto s = ctx->get_saved_variables();
    Tensor gn = virial_grad_impl(g[0].contiguous(), s[1], s[2], s[3],
                                 ctx->saved_data["nloc"].toInt(), s[0].scalar_type());
    return {gn.view(s[0].sizes()), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
  }
};

std::tuple<Tensor, Tensor, Tensor> prod_env_mat_a(const Tensor& coord, const Tensor& atype,
                                                  const Tensor& nlist, const Tensor& davg,
                                                  const Tensor& dstd, double rcut,
                                                  double rcut_smth, double protection) {
  auto r = ProdEnvMatAOp::apply(coord, atype, nlist, davg, dstd, rcut, rcut_smth, protection);
  return {r[0], r[1], r[2]};
}

Tensor prod_force_se_a(const Tensor& net_deriv, const Tensor& in_deriv, const Tensor& nlist,
                       int64_t nloc, int64_t nall) {
  return ProdForceSeAOp::apply(net_deriv, in_deriv, nlist, nloc, nall)[0];
}

std::tuple<Tensor, Tensor> prod_virial_se_a(const Tensor& net_deriv, const Tensor& in_deriv,
                                            const Tensor& rij, const Tensor& nlist, int64_t nloc,
                                            int64_t nall) {
  auto r = ProdVirialSeAOp::apply(net_deriv, in_deriv, rij, nlist, nloc, nall);
  return {r[0], r[1]};
}

Tensor prod_force_se_a_grad(const Tensor& grad, const Tensor& net_deriv, const Tensor& in_deriv,
                            const Tensor& nlist, int64_t nloc) {
  return force_grad_impl(grad, in_deriv, nlist, nloc, nloc, net_deriv.scalar_type());
}

Tensor prod_virial_se_a_grad(const Tensor& grad, const Tensor& net_deriv, const Tensor& in_deriv,
                             const Tensor& rij, const Tensor& nlist, int64_t nloc) {
  return virial_grad_impl(grad, in_deriv, rij, nlist, nloc, net_deriv.scalar_type());
}

Tensor enable_mpi() { return torch::zeros({1}, torch::kBool); }

}  // namespace

TORCH_LIBRARY(deepmd, m) { m.def("enable_mpi", enable_mpi); }

TORCH_LIBRARY_FRAGMENT(deepmd, m) {
  // schema inferred from the C++ signature, as the reference registers it
  m.def("tabulate_fusion_se_a", tabulate_fusion_se_a);
  m.def("tabulate_fusion_se_a_grad(Tensor table, Tensor table_info, Tensor em_x, Tensor em, "
        "Tensor dy, Tensor descriptor) -> Tensor[]",
        tabulate_fusion_se_a_grad);
  m.def("tabulate_fusion_se_a_grad_grad(Tensor table, Tensor table_info, Tensor em_x, "
        "Tensor em, Tensor dz_dy_dem_x, Tensor dz_dy_dem, Tensor descriptor) -> Tensor",
        tabulate_fusion_se_a_grad_grad);
  m.def("prod_env_mat_a(Tensor coord, Tensor atype, Tensor nlist, Tensor davg, Tensor dstd, "
        "float rcut, float rcut_smth, float protection=0.0) -> (Tensor, Tensor, Tensor)",
        prod_env_mat_a);
  m.def("prod_force_se_a(Tensor net_deriv, Tensor in_deriv, Tensor nlist, int nloc, int nall) "
        "-> Tensor",
        prod_force_se_a);
  m.def("prod_virial_se_a(Tensor net_deriv, Tensor in_deriv, Tensor rij, Tensor nlist, "
        "int nloc, int nall) -> (Tensor, Tensor)",
        prod_virial_se_a);
  m.def("prod_force_se_a_grad(Tensor grad, Tensor net_deriv, Tensor in_deriv, Tensor nlist, "
        "int nloc) -> Tensor",
        prod_force_se_a_grad);
  m.def("prod_virial_se_a_grad(Tensor grad, Tensor net_deriv, Tensor in_deriv, Tensor rij, "
        "Tensor nlist, int nloc) -> Tensor",
        prod_virial_se_a_grad);
}
