"""
ChainRulesCore `rrule` for `logabspfaffian` (src/pfaffian.jl:235-236 for SkewHermitian,
:244-257 for a dense skew-symmetric matrix) on the GPU: the forward log|Pf| and sign and the
cotangent `Δlogpf · (-½A⁻¹)` come from one `skh_logpf_grad_batched_d` call.  The sign output is
piecewise constant, so its cotangent contributes nothing.  The reference has no AD code; its CPU
counterpart of the gradient is `-½ inv(A)` with `inv(A::SkewHermitian) = skewhermitian!(inv(A.data))`
(src/skewhermitian.jl:233).
"""
module SkewLinearAlgebraHIPChainRulesCoreExt

using ChainRulesCore
using LinearAlgebra
using SkewLinearAlgebra
using SkewLinearAlgebra: SkewHermitian
using SkewLinearAlgebraHIP: logabspfaffian_with_gradient

# cotangent of the log|Pf| output; the sign's is dropped
function _dlogpf(Δ)
    Δ isa AbstractZero && return Δ
    return unthunk(unthunk(Δ)[1])
end

# A::SkewHermitian{Float64}: a structural tangent for the wrapped matrix
function ChainRulesCore.rrule(::typeof(SkewLinearAlgebra.logabspfaffian),
                              A::SkewHermitian{Float64,Matrix{Float64}})
    lp, sg, G = logabspfaffian_with_gradient(A)
    function logabspfaffian_pullback(Δ)
        d = _dlogpf(Δ)
        d isa AbstractZero && return NoTangent(), ZeroTangent()
        return NoTangent(), Tangent{typeof(A)}(; data = d .* G)
    end
    return (lp, sg), logabspfaffian_pullback
end

# a dense skew-symmetric Matrix{Float64}: the cotangent is the matrix d · (-½A⁻¹)
function ChainRulesCore.rrule(::typeof(SkewLinearAlgebra.logabspfaffian), A::Matrix{Float64})
    lp, sg, G = logabspfaffian_with_gradient(A)
    function logabspfaffian_pullback(Δ)
        d = _dlogpf(Δ)
        d isa AbstractZero && return NoTangent(), ZeroTangent()
        return NoTangent(), d .* G
    end
    return (lp, sg), logabspfaffian_pullback
end

end # module
