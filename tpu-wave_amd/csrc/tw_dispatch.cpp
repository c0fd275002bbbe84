// tw_dispatch.cpp -- routes the precision-dispatched kernel entry points of
// tw_internal.h to the build of tw_kernels.hip for the plan's storage type
// (tw_kernels.hip is compiled once per type with -DTW_PREC=0..3 and suffixes
// these names with _p0.._p3).
#include "tw_internal.h"

namespace tw {

#define TW_DECL(p)                                                                                                  \
  size_t stat_fill_p##p(const Plan&, void*, int);                                                                   \
  hipError_t launch_half_p##p(const Plan&, int, const StepCtl&, hipStream_t);                                      \
  hipError_t launch_stress_energy_p##p(const Plan&, hipStream_t);                                                  \
  hipError_t launch_uhat_p##p(const Plan&, hipStream_t);                                                           \
  hipError_t launch_finalize_p##p(const Plan&, const Pending&, hipStream_t);                                       \
  hipError_t launch_fused_p##p(const Plan&, int, const StepCtl&, hipStream_t);                                     \
  hipError_t launch_batch_half_p##p(const Plan&, int, const BatchCore*, int, int, const BatchBand*, int, int,       \
                                    hipStream_t);                                                                   \
  hipError_t launch_adjoint_step_p##p(const Plan&, const AdjPlan&, const StateView&, const StateView&, int, int,   \
                                      hipStream_t);
TW_DECL(0)
TW_DECL(1)
TW_DECL(2)
TW_DECL(3)

#define TW_ROUTE(name, ...)                     \
  switch (p.precision) {                        \
    case TW_FP64: return name##_p0(__VA_ARGS__); \
    case TW_FP32: return name##_p1(__VA_ARGS__); \
    case TW_BF16: return name##_p2(__VA_ARGS__); \
    default: return name##_p3(__VA_ARGS__);      \
  }

size_t stat_fill(const Plan& p, void* out, int pp) { TW_ROUTE(stat_fill, p, out, pp) }
hipError_t launch_half(const Plan& p, int half, const StepCtl& c, hipStream_t s) { TW_ROUTE(launch_half, p, half, c, s) }
hipError_t launch_stress_energy(const Plan& p, hipStream_t s) { TW_ROUTE(launch_stress_energy, p, s) }
hipError_t launch_uhat(const Plan& p, hipStream_t s) { TW_ROUTE(launch_uhat, p, s) }
hipError_t launch_finalize(const Plan& p, const Pending& pend, hipStream_t s) { TW_ROUTE(launch_finalize, p, pend, s) }
hipError_t launch_fused(const Plan& p, int what, const StepCtl& c, hipStream_t s) {
  TW_ROUTE(launch_fused, p, what, c, s)
}
hipError_t launch_batch_half(const Plan& p, int half, const BatchCore* core, int ncore, int core_blocks,
                             const BatchBand* band, int nband, int band_blocks, hipStream_t s) {
  TW_ROUTE(launch_batch_half, p, half, core, ncore, core_blocks, band, nband, band_blocks, s)
}
hipError_t launch_adjoint_step(const Plan& p, const AdjPlan& a, const StateView& cur, const StateView& prev, int it,
                               int it_rel, hipStream_t s) {
  TW_ROUTE(launch_adjoint_step, p, a, cur, prev, it, it_rel, s)
}

}  // namespace tw
