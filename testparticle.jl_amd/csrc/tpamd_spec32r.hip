// tpamd_spec32r.hip -- the specialised gridded-B FP32C kernels of the
// reference-form multistep update (float storage, node or cell layout).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_spec32r(const LaunchReq& r, int emode, int ns, bool per, bool cells)
{
    with_launcher(r, [&](const auto& l) {
        using L = std::decay_t<decltype(l)>;
        if (cells) {
            if (emode == kEUxB) sel_g32_nsr<L, kEUxB, true>(l, ns, per);
            else sel_g32_nsr<L, kEUniform, true>(l, ns, per);
        } else {
            if (emode == kEUxB) sel_g32_nsr<L, kEUxB, false>(l, ns, per);
            else sel_g32_nsr<L, kEUniform, false>(l, ns, per);
        }
    });
}

}  // namespace tpamd
