"""
    SkewLinearAlgebraHIP

Runs SkewLinearAlgebra.jl's skew tridiagonalization path on an MI355X through `libskewhip.so`
(the C ABI of `include/skewhip.h`).  Loading this package replaces the seam every consumer of the
path calls, `skewblockedhess!(S::SkewHermitian{T}) -> (τ, η)` (src/hessenberg.jl:144-174), for
`Matrix{Float64}` and `Matrix{ComplexF64}` storage; `hessenberg!` (src/hessenberg.jl:5-13),
`eigvals!`/`eigen!` (src/eigen.jl:49-97), `_pfaffian!` and `_logabspfaffian!`
(src/pfaffian.jl:74-94,167-179) then run on the GPU unchanged, because the factors, τ and η
come back in the reference's layout.  Other storage types keep the reference's CPU methods.

Also: the host implicit-QR `skewtrieigvals!` (src/skeweigen.jl:105-156), batched log|Pf| and
the log|Pf| gradient `-½A⁻¹` (no reference counterpart; the CPU route is `inv(A)`,
src/skewhermitian.jl:233).  The ChainRulesCore `rrule` for `logabspfaffian` is the package
extension `ext/SkewLinearAlgebraHIPChainRulesCoreExt.jl`.

The library path is `ENV["SKEWHIP_LIB"]` at load time, else `libskewhip.so` on the loader path.
"""
module SkewLinearAlgebraHIP

using LinearAlgebra
using SkewLinearAlgebra
using SkewLinearAlgebra: SkewHermitian, SkewHermTridiagonal

export logabspfaffian_batched, logabspfaffian_with_gradient, logabspfaffian_grad_batched

const libskewhip = get(ENV, "SKEWHIP_LIB", "libskewhip.so")

# skh_opts of include/skewhip.h: panel width (0 = the library's table) and engine flags (0 = auto)
struct SkhOpts
    nb::Int64
    flags::Int64
end

const SKH_ERR_NOCONV = 2

function _check(rc::Integer, name::AbstractString)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:skh_last_error, libskewhip), Cstring, ()))
    rc < 0 && throw(ArgumentError("$name: invalid argument $(-rc): $msg"))
    rc == SKH_ERR_NOCONV && error("Maximum number of iterations reached, the algorithm didn't converge")
    error("$name failed (status $rc): $msg")
end

# ---- the seam: skewblockedhess! (src/hessenberg.jl:144-174) -----------------------------------

# Float64: A.data becomes the 'L'-layout factors (unit stored at the head of each reflector);
# τ, η as the reference returns them.
function SkewLinearAlgebra.skewblockedhess!(S::SkewHermitian{Float64,Matrix{Float64}})
    A = S.data
    n = LinearAlgebra.checksquare(A)
    τ = Vector{Float64}(undef, max(n - 1, 0))
    η = Vector{Float64}(undef, max(n - 1, 0))
    nz = Ref{Int64}(0)
    opts = Ref(SkhOpts(0, 0))
    rc = GC.@preserve A τ η ccall((:skh_sytrd_d, libskewhip), Cint,
                                  (Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Ref{Int64}, Ref{SkhOpts}),
                                  n, A, max(stride(A, 2), 1), τ, η, nz, opts)
    _check(rc, "skh_sytrd_d")
    return τ, η
end

# ComplexF64: the complex branch (src/hessenberg.jl:92-131 with :96-104, :123-126, :161, tail
# :70-90).  The library leaves T's imaginary diagonal on A.data's diagonal, where hessenberg!
# reads it (`imag(diag(A.data))`, src/hessenberg.jl:7-8); η is real, τ complex.
function SkewLinearAlgebra.skewblockedhess!(S::SkewHermitian{ComplexF64,Matrix{ComplexF64}})
    A = S.data
    n = LinearAlgebra.checksquare(A)
    τ = Vector{ComplexF64}(undef, max(n - 1, 0))
    η = Vector{Float64}(undef, max(n - 1, 0))
    dvim = Vector{Float64}(undef, max(n, 1))
    opts = Ref(SkhOpts(0, 0))
    rc = GC.@preserve A τ η dvim ccall((:skh_sytrd_z, libskewhip), Cint,
                                       (Int64, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Ptr{Float64}, Ptr{Float64},
                                        Ref{SkhOpts}),
                                       n, A, max(stride(A, 2), 1), τ, η, dvim, opts)
    _check(rc, "skh_sytrd_z")
    return τ, η
end

# ---- host implicit QR on the tridiagonal (src/skeweigen.jl:105-156, backup :463-468) -----------
function SkewLinearAlgebra.skewtrieigvals!(T::SkewHermTridiagonal{Float64,Vector{Float64},Nothing})
    n = size(T, 1)
    vals = Vector{Float64}(undef, n)
    ev = T.ev
    rc = GC.@preserve ev vals ccall((:skh_skewtrieigvals_d, libskewhip), Cint, (Int64, Ptr{Float64}, Ptr{Float64}),
                                    n, ev, vals)
    _check(rc, "skh_skewtrieigvals_d")
    return complex.(0, vals)
end

# ---- batched log|Pf| and the gradient (no reference counterparts) ------------------------------

# matrices packed column-major one after another, each at an even element offset
function _pack(mats::AbstractVector{<:AbstractMatrix{Float64}})
    ns = Int64[LinearAlgebra.checksquare(M) for M in mats]
    offs = Vector{Int64}(undef, length(mats))
    t = 0
    for b in eachindex(mats)
        offs[b] = t
        t += ns[b]^2 + isodd(ns[b]^2)
    end
    packed = zeros(Float64, max(t, 1))
    for b in eachindex(mats)
        copyto!(packed, offs[b] + 1, mats[b], 1, ns[b]^2)
    end
    return ns, offs, packed
end

"""
    logabspfaffian_batched(mats) -> (logpf, sgn)

`logabspfaffian` (src/pfaffian.jl:167-179, 235-236) of every real skew matrix in `mats` (uneven
sizes allowed) in one persistent GPU launch.  Odd n gives (-Inf, 0), as the reference.
"""
function logabspfaffian_batched(mats::AbstractVector{<:AbstractMatrix{Float64}})
    ns, offs, packed = _pack(mats)
    lp = Vector{Float64}(undef, length(mats))
    sg = similar(lp)
    rc = GC.@preserve ns offs packed lp sg ccall((:skh_logpf_batched_d, libskewhip), Cint,
                                                 (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64},
                                                  Ptr{Float64}, Ptr{SkhOpts}),
                                                 length(mats), ns, offs, packed, lp, sg, C_NULL)
    _check(rc, "skh_logpf_batched_d")
    return lp, sg
end

"""
    logabspfaffian_grad_batched(mats) -> (logpf, sgn, grads)

log|Pf|, sign and `∂ log|Pf(A)| / ∂A = -½A⁻¹` (entrywise convention) of every real skew matrix in
`mats`; A⁻¹ = Q T⁻¹ Qᵀ from the stored reflectors with the closed-form inverse of the skew
tridiagonal T.  The CPU route of the same quantity is `-½ inv(A)` (src/skewhermitian.jl:233).
Odd n: (-Inf, 0) and a zero gradient.
"""
function logabspfaffian_grad_batched(mats::AbstractVector{<:AbstractMatrix{Float64}})
    ns, offs, packed = _pack(mats)
    lp = Vector{Float64}(undef, length(mats))
    sg = similar(lp)
    g = similar(packed)
    rc = GC.@preserve ns offs packed lp sg g ccall((:skh_logpf_grad_batched_d, libskewhip), Cint,
                                                   (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64},
                                                    Ptr{Float64}, Ptr{Float64}, Ptr{SkhOpts}),
                                                   length(mats), ns, offs, packed, lp, sg, g, C_NULL)
    _check(rc, "skh_logpf_grad_batched_d")
    grads = [reshape(g[offs[b]+1:offs[b]+ns[b]^2], ns[b], ns[b]) for b in eachindex(mats)]
    return lp, sg, grads
end

"""
    logabspfaffian_with_gradient(A) -> (logpf, sgn, G)

`logabspfaffian(A)` and `G = ∂ log|Pf(A)| / ∂A = -½A⁻¹` for one real skew matrix (a
`SkewHermitian{Float64}` or a skew-symmetric `Matrix{Float64}`, checked as the reference's
`logabspfaffian!` does, src/pfaffian.jl:254-257).
"""
function logabspfaffian_with_gradient(A::AbstractMatrix{Float64})
    D = A isa SkewHermitian ? A.data : A
    A isa SkewHermitian || isskewhermitian(D) ||
        throw(ArgumentError("Pfaffian requires a skew-Hermitian matrix"))
    lp, sg, gs = logabspfaffian_grad_batched([Matrix{Float64}(D)])
    return lp[1], sg[1], gs[1]
end

end # module
