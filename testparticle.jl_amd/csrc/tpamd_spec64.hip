// tpamd_spec64.hip -- the specialised gridded-B FP64 kernels: B a uniform
// grid (double or float storage, node or cell layout: one line per gather
// for float cells, two for double cells), E zero, uniform or UxB, the Boris
// update or the rational multistep update with n <= 2.
#include "tpamd_pusher.h"

namespace tpamd {

void launch_spec64(const LaunchReq& r, int emode, int ns, bool per, bool f32, bool cells)
{
    with_launcher(r, [&](const auto& l) {
        using L = std::decay_t<decltype(l)>;
        if (f32 && cells) {
            if (emode == kEUxB) sel_g64_ns<L, kEUxB, float, true>(l, ns, per);
            else sel_g64_ns<L, kEUniform, float, true>(l, ns, per);
        } else if (f32) {
            if (emode == kEUxB) sel_g64_ns<L, kEUxB, float, false>(l, ns, per);
            else sel_g64_ns<L, kEUniform, float, false>(l, ns, per);
        } else if (cells) {
            if (emode == kEUxB) sel_g64_ns<L, kEUxB, double, true>(l, ns, per);
            else sel_g64_ns<L, kEUniform, double, true>(l, ns, per);
        } else {
            if (emode == kEUxB) sel_g64_ns<L, kEUxB, double, false>(l, ns, per);
            else sel_g64_ns<L, kEUniform, double, false>(l, ns, per);
        }
    });
}

}  // namespace tpamd
