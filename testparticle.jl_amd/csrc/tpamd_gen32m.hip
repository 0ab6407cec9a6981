// tpamd_gen32m.hip -- the generic FP32C kernels that also evaluate mesh fields (through the
// out-of-line mesh evaluator).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_generic32_mesh(const LaunchReq& r, int alg)
{
    with_launcher(r, [&](const auto& l) {
        if (alg == 0) l.template go<S32CM, 0>();
        else l.template go<S32CM, 1>();
    });
}

}  // namespace tpamd
