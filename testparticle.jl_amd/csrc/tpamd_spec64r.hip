// tpamd_spec64r.hip -- the specialised gridded-B FP64 kernels of the
// reference-form multistep update (atan, sincos and the Taylor branch of
// src/boris/multistep_boris.jl:94-120, n = 1 or 2), in a unit of their own so
// that the kernel families compile in parallel.
#include "tpamd_pusher.h"

namespace tpamd {

void launch_spec64r(const LaunchReq& r, int emode, int ns, bool per, bool f32, bool cells)
{
    with_launcher(r, [&](const auto& l) {
        using L = std::decay_t<decltype(l)>;
        if (f32 && cells) {
            if (emode == kEUxB) sel_g64_nsr<L, kEUxB, float, true>(l, ns, per);
            else sel_g64_nsr<L, kEUniform, float, true>(l, ns, per);
        } else if (f32) {
            if (emode == kEUxB) sel_g64_nsr<L, kEUxB, float, false>(l, ns, per);
            else sel_g64_nsr<L, kEUniform, float, false>(l, ns, per);
        } else if (cells) {
            if (emode == kEUxB) sel_g64_nsr<L, kEUxB, double, true>(l, ns, per);
            else sel_g64_nsr<L, kEUniform, double, true>(l, ns, per);
        } else {
            if (emode == kEUxB) sel_g64_nsr<L, kEUxB, double, false>(l, ns, per);
            else sel_g64_nsr<L, kEUniform, double, false>(l, ns, per);
        }
    });
}

}  // namespace tpamd
