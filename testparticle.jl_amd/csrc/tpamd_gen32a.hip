// tpamd_gen32a.hip -- the generic FP32C kernels for analytic fields (zero, uniform, affine,
// dipole, E = -u x B of those).
#include "tpamd_pusher.h"

namespace tpamd {

void launch_generic32_analytic(const LaunchReq& r, int alg, bool special)
{
    // fixed-step Boris in fields that are the same everywhere: with no
    // boundary, one particle per quad of lanes (four threads a particle);
    // else the kernel that forms the update's rotation vectors once per
    // particle
    if (special && !r.batch && !r.adaptive && alg == 0 && r.p->uniform_fields && r.p->rec_w == 6 &&
        r.p->bnd_flags == 0) {
        const long long threads = 4 * r.p->N;
        const dim3 g((unsigned)((threads + r.b.x - 1) / r.b.x));
        hipLaunchKernelGGL(k_fixed_quad32, g, r.b, 0, r.st, *r.p);
        return;
    }
    if (special && !r.batch && !r.adaptive && alg == 0 && r.p->uniform_fields && r.p->rec_w == 6) {
        SoloLaunch{*r.p, 0, r.g, r.b, r.st}.template go<Hoisted<S32CA>, 0>();
        return;
    }
    // fixed-step Boris in a dipole B with no E: the kernel compiled for
    // that pair of fields
    if (special && !r.batch && !r.adaptive && alg == 0 && r.p->rec_w == 6 && r.p->B.kind == kFieldDipole &&
        r.p->E.kind == kFieldZero && !r.p->e_is_uxb) {
        hipLaunchKernelGGL((k_fixed<DipoleOnly<S32CA>, 0>), r.g, r.b, 0, r.st, *r.p);
        return;
    }
    with_launcher(r, [&](const auto& l) {
        if (alg == 0) l.template go<S32CA, 0>();
        else l.template go<S32CA, 1>();
    });
}

}  // namespace tpamd
