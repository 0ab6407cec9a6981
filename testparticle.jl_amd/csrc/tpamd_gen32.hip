// tpamd_gen32.hip -- the generic FP32C kernels for analytic and gridded fields
// (uniform grids and their time sequences).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_generic32(const LaunchReq& r, int alg)
{
    with_launcher(r, [&](const auto& l) {
        if (alg == 0) l.template go<S32C, 0>();
        else l.template go<S32C, 1>();
    });
}

}  // namespace tpamd
