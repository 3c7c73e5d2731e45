// Thin PyTorch-ROCm extension over the C ABI: the apply of a prepared plan as a registered
// operator, torch.ops.skelfmm.apply(plan, x, y), launched on the caller's current HIP stream
// (reference: SkelFMMBatched.matvec, skelfmm/batched_fmm.py:539-564). The plan comes from
// skf_plan_create; the operator checks the tensors against each other and hands their device
// pointers to skf_apply, which runs the whole apply (gather, upward levels, near field,
// downward levels, scatter) without allocating or synchronizing.
#include <ATen/core/Tensor.h>
#include <ATen/hip/impl/HIPGuardImplMasqueradingAsCUDA.h>
#include <ATen/hip/impl/HIPStreamMasqueradingAsCUDA.h>
#include <torch/library.h>

#include "skelfmm_hip.h"

#ifndef SKF_SOURCE_HASH
#define SKF_SOURCE_HASH "0000000000000000"
#endif
// found by the loader without loading the library (same marker as in the HIP library)
extern "C" __attribute__((used)) const char skf_torch_build_tag[] = "skelfmm-build-id:" SKF_SOURCE_HASH;

namespace {

void apply(int64_t plan, const at::Tensor& x, const at::Tensor& y) {
  TORCH_CHECK(plan != 0, "skelfmm::apply: null plan handle");
  TORCH_CHECK(x.device().is_cuda() && y.device() == x.device(), "skelfmm::apply: x and y must be on one GPU");
  TORCH_CHECK(x.scalar_type() == y.scalar_type() &&
                  (x.scalar_type() == at::kFloat || x.scalar_type() == at::kDouble),
              "skelfmm::apply: x and y must both be float32 or float64 (real views of complex vectors)");
  TORCH_CHECK(x.numel() == y.numel(), "skelfmm::apply: x and y must have the same size");
  TORCH_CHECK(x.is_contiguous() && y.is_contiguous(), "skelfmm::apply: x and y must be contiguous");
  // PyTorch-ROCm names GPU devices "cuda": the guard and stream types that accept them
  const c10::hip::HIPGuardMasqueradingAsCUDA guard(x.device());
  hipStream_t stream = c10::hip::getCurrentHIPStreamMasqueradingAsCUDA(x.device().index()).stream();
  const int rc = skf_apply(reinterpret_cast<skf_plan*>(plan), x.data_ptr(), y.data_ptr(), stream);
  TORCH_CHECK(rc == 0, "skelfmm::apply: ", skf_last_error());
}

}  // namespace

TORCH_LIBRARY(skelfmm, m) {
  m.def("apply(int plan, Tensor x, Tensor(a!) y) -> ()");
}

TORCH_LIBRARY_IMPL(skelfmm, CUDA, m) {
  m.impl("apply", &apply);
}
