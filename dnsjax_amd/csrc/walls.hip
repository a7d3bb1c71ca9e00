// Wall-bounded (plane Couette) path: placeholder until the IMM kernels land.
#include <hip/hip_runtime.h>

#include "context.hpp"

namespace dnsmi {
int walls_setup(Ctx* c) { return c->fail(DNSMI_EINVAL, "plane Couette not built yet"); }
void walls_release(Ctx*) {}
int walls_get_rhs(Ctx* c, const double*, double*, hipStream_t) { return c->fail(DNSMI_EINVAL, "walls"); }
int walls_predict(Ctx* c, const double*, const double*, double*, hipStream_t) { return c->fail(DNSMI_EINVAL, "walls"); }
int walls_correct(Ctx* c, const double*, const double*, const double*, const double*, double*, double*, hipStream_t) { return c->fail(DNSMI_EINVAL, "walls"); }
int walls_norm(Ctx* c, const double*, double*, hipStream_t) { return c->fail(DNSMI_EINVAL, "walls"); }
int walls_predict_and_correct(Ctx* c, const double*, double*, double*, double*, hipStream_t) { return c->fail(DNSMI_EINVAL, "walls"); }
int walls_iterate_correction(Ctx* c, const double*, double*, const double*, double*, double*, hipStream_t) { return c->fail(DNSMI_EINVAL, "walls"); }
}  // namespace dnsmi
