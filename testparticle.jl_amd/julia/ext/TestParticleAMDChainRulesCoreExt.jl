# TestParticleAMDChainRulesCoreExt.jl -- the reverse rule of
# TestParticleAMD.trajectory_loss, loaded with ChainRulesCore.
#
# The loss and both gradients come from one tpamd_adjoint call (a forward pass
# with FP64 checkpoints, then the reverse sweep), so the pullback only scales
# them by the incoming cotangent.  The reference's counterpart is ForwardDiff
# through its CPU solve (test/test_derivatives.jl); Zygote and other
# ChainRules-based reverse-mode tools pick this rule up instead.
module TestParticleAMDChainRulesCoreExt

using TestParticleAMD
using ChainRulesCore

function ChainRulesCore.rrule(::typeof(TestParticleAMD.trajectory_loss), A::Array{<:Real, 4},
        xv0::AbstractMatrix, targets::AbstractMatrix; kw...)
    loss, gA, gx = TestParticleAMD.loss_gradient(A, xv0, targets; kw...)
    function trajectory_loss_pullback(dl)
        s = unthunk(dl)
        # cotangents shaped like the primal arguments: A's 3 x nx x ny x nz
        # array and xv0's 6 x N matrix; the targets get none
        return NoTangent(), ProjectTo(A)(s .* gA), ProjectTo(xv0)(s .* gx), NoTangent()
    end
    return loss, trajectory_loss_pullback
end

end # module
