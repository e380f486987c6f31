// Fused ZDT path for power-of-two Nf.  (Not yet enabled: fast_supported
// returns false and the generic kernels serve every shape.)
#include <hip/hip_runtime.h>

#include "zdt_fast.h"

namespace zdt {

struct FastPlan {
  int64_t Nf = 0;
};

bool fast_supported(int64_t, int64_t, int64_t) { return false; }
int64_t fast_bytes(int64_t, int64_t, int64_t) { return 0; }
int fast_create(int64_t, int64_t, int64_t, float, FastPlan**) {
  return fail(ZDT_ENOTSUP, "fused path not available");
}
void fast_destroy(FastPlan* p) { delete p; }
int fast_input(FastPlan*, const float*, hipStream_t) { return fail(ZDT_ENOTSUP, "fused path not available"); }
int fast_unpack_F(FastPlan*, c32*, hipStream_t) { return fail(ZDT_ENOTSUP, "fused path not available"); }
int fast_run(FastPlan*, float, int, const float*, float*, double, hipStream_t) {
  return fail(ZDT_ENOTSUP, "fused path not available");
}
int fast_output(FastPlan*, float*, double, hipStream_t) { return fail(ZDT_ENOTSUP, "fused path not available"); }

}  // namespace zdt
