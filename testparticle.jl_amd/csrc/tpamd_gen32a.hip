// tpamd_gen32a.hip -- the generic FP32C kernels for analytic fields (zero, uniform, affine,
// dipole, E = -u x B of those).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_generic32_analytic(const LaunchReq& r, int alg)
{
    with_launcher(r, [&](const auto& l) {
        if (alg == 0) l.template go<S32CA, 0>();
        else l.template go<S32CA, 1>();
    });
}

}  // namespace tpamd
