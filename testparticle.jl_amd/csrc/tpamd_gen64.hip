// tpamd_gen64.hip -- the generic FP64 kernels for analytic and gridded fields
// (uniform grids and their time sequences).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_generic64(const LaunchReq& r, int alg)
{
    with_launcher(r, [&](const auto& l) {
        if (alg == 0) l.template go<S64, 0>();
        else l.template go<S64, 1>();
    });
}

}  // namespace tpamd
