// torch::autograd::Function around the two-site H_eff matvec: the forward is
// dmrgx_heff_apply, the backward dmrgx_heff_adjoint (cotangents of psi and
// of both environments).  Torch owns the tensors; the library runs on its own
// stream, so each call synchronises the device on entry and on exit.
#include <torch/cuda.h>
#include <torch/extension.h>

#include <stdexcept>
#include <string>

#include "dmrgx.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void ok(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + dmrgx_last_error());
}

// The get_map structure of one two-site step: maps (int32, 7 per task),
// block offsets / sizes of each environment (int64), all on the host.
struct Structure {
  Tensor rmap, roff, rsize, lmap, loff, lsize;
  int64_t n_store;
  int prec;
};

dmrgx_heff* make_plan(const Structure& st, const Tensor& renv, const Tensor& lenv, int64_t n_psi) {
  const int dt = renv.is_complex() ? DMRGX_C64 : DMRGX_R64;
  dmrgx_heff* h = nullptr;
  ok(dmrgx_heff_create(&h, dt, st.prec, renv.data_ptr(), st.roff.data_ptr<int64_t>(), st.rsize.data_ptr<int64_t>(),
                       (int)st.roff.numel(), st.rmap.data_ptr<int32_t>(), (int)(st.rmap.numel() / 7), lenv.data_ptr(),
                       st.loff.data_ptr<int64_t>(), st.lsize.data_ptr<int64_t>(), (int)st.loff.numel(),
                       st.lmap.data_ptr<int32_t>(), (int)(st.lmap.numel() / 7), n_psi, st.n_store),
     "dmrgx_heff_create");
  return h;
}

void check_operand(const Tensor& t, const Tensor& like, const char* name) {
  TORCH_CHECK(t.is_cuda(), name, " must be a device tensor");
  TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
  TORCH_CHECK(t.scalar_type() == like.scalar_type(), name, " must have psi's dtype");
  TORCH_CHECK(t.scalar_type() == torch::kDouble || t.scalar_type() == torch::kComplexDouble,
              name, " must be float64 or complex128");
}

class HeffFunction : public torch::autograd::Function<HeffFunction> {
 public:
  static Tensor forward(AutogradContext* ctx, Tensor psi, Tensor renv, Tensor lenv, Tensor rmap, Tensor roff,
                        Tensor rsize, Tensor lmap, Tensor loff, Tensor lsize, int64_t n_store, int64_t prec) {
    psi = psi.contiguous();
    renv = renv.contiguous();
    lenv = lenv.contiguous();
    check_operand(psi, psi, "psi");
    check_operand(renv, psi, "renv");
    check_operand(lenv, psi, "lenv");
    Structure st{rmap.contiguous().cpu().to(torch::kInt32), roff.contiguous().cpu().to(torch::kInt64),
                 rsize.contiguous().cpu().to(torch::kInt64), lmap.contiguous().cpu().to(torch::kInt32),
                 loff.contiguous().cpu().to(torch::kInt64), lsize.contiguous().cpu().to(torch::kInt64), n_store,
                 (int)prec};
    ctx->save_for_backward({psi, renv, lenv});
    ctx->saved_data["rmap"] = st.rmap;
    ctx->saved_data["roff"] = st.roff;
    ctx->saved_data["rsize"] = st.rsize;
    ctx->saved_data["lmap"] = st.lmap;
    ctx->saved_data["loff"] = st.loff;
    ctx->saved_data["lsize"] = st.lsize;
    ctx->saved_data["n_store"] = n_store;
    ctx->saved_data["prec"] = prec;
    Tensor out = torch::empty_like(psi);
    torch::cuda::synchronize();
    dmrgx_heff* h = make_plan(st, renv, lenv, psi.numel());
    const int rc = dmrgx_heff_apply(h, psi.data_ptr(), out.data_ptr());
    dmrgx_heff_destroy(h);
    ok(rc, "dmrgx_heff_apply");
    return out;
  }

  static variable_list backward(AutogradContext* ctx, variable_list grads) {
    auto saved = ctx->get_saved_variables();
    Tensor psi = saved[0], renv = saved[1], lenv = saved[2];
    Structure st{ctx->saved_data["rmap"].toTensor(), ctx->saved_data["roff"].toTensor(),
                 ctx->saved_data["rsize"].toTensor(), ctx->saved_data["lmap"].toTensor(),
                 ctx->saved_data["loff"].toTensor(), ctx->saved_data["lsize"].toTensor(),
                 ctx->saved_data["n_store"].toInt(), (int)ctx->saved_data["prec"].toInt()};
    Tensor g = grads[0].contiguous();
    const bool want_p = ctx->needs_input_grad(0), want_r = ctx->needs_input_grad(1),
               want_l = ctx->needs_input_grad(2);
    Tensor gp = want_p ? torch::empty_like(psi) : Tensor();
    Tensor gr = want_r ? torch::empty_like(renv) : Tensor();
    Tensor gl = want_l ? torch::empty_like(lenv) : Tensor();
    torch::cuda::synchronize();
    dmrgx_heff* h = make_plan(st, renv, lenv, psi.numel());
    const int rc = dmrgx_heff_adjoint(h, psi.data_ptr(), g.data_ptr(), want_p ? gp.data_ptr() : nullptr,
                                      want_r ? gr.data_ptr() : nullptr, want_l ? gl.data_ptr() : nullptr);
    dmrgx_heff_destroy(h);
    ok(rc, "dmrgx_heff_adjoint");
    ok(dmrgx_synchronize(), "dmrgx_synchronize");
    return {gp, gr, gl, Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
  }
};

Tensor heff_apply(Tensor psi, Tensor renv, Tensor lenv, Tensor rmap, Tensor roff, Tensor rsize, Tensor lmap,
                  Tensor loff, Tensor lsize, int64_t n_store, int64_t prec) {
  return HeffFunction::apply(psi, renv, lenv, rmap, roff, rsize, lmap, loff, lsize, n_store, prec);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.def("heff_apply", &heff_apply,
        "H_eff psi with autograd: (psi, renv, lenv, rmap, roff, rsize, lmap, loff, lsize, n_store, prec)");
  m.def("init", [](int device) { ok(dmrgx_init(device), "dmrgx_init"); });
}
