// tpamd_gen64m.hip -- the generic FP64 kernels that also evaluate mesh fields (through the
// out-of-line mesh evaluator).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_generic64_mesh(const LaunchReq& r, int alg)
{
    with_launcher(r, [&](const auto& l) {
        if (alg == 0) l.template go<S64M, 0>();
        else l.template go<S64M, 1>();
    });
}

}  // namespace tpamd
