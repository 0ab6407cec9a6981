// tpamd_gen64a.hip -- the generic FP64 kernels for analytic fields (zero, uniform, affine,
// dipole, E = -u x B of those).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_generic64_analytic(const LaunchReq& r, int alg)
{
    with_launcher(r, [&](const auto& l) {
        if (alg == 0) l.template go<S64A, 0>();
        else l.template go<S64A, 1>();
    });
}

}  // namespace tpamd
