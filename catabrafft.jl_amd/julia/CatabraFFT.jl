# CatabraFFT.jl drop-in: the reference module's user API (CatabraFFT.jl:39-226, spells.jl:4-37)
# implemented by `ccall`s into the MI355X library lib/libcatabrafft.so (include/catabrafft.h).
#
#   fft / ifft / bfft      -> cfft_exec_host on a cached (n, T, NO_FLAG) plan
#   plan_fft(x, flags)     -> cfft_plan_1d(..., flags): the returned Spell's flag is honoured
#   mul!(y, p, x), p * x   -> cfft_exec_host with p's plan (x is never overwritten)
#   empty_cache()          -> cfft_empty_cache + dropping the plan handles
# Device arrays (AMDGPU.jl ROCArray) go through exec_device!, i.e. cfft_exec with device pointers.
module CatabraFFT

using AbstractFFTs
import LinearAlgebra: mul!

const libcfft = joinpath(@__DIR__, "..", "lib", "libcatabrafft.so")

const FLAG = Int8
const NO_FLAG = FLAG(0)
const MEASURE = FLAG(1)
const ENCHANT = FLAG(2)

const FORWARD = Cint(-1)
const BACKWARD = Cint(1)

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
const LOCK = ReentrantLock()

function check(rc::Cint)
    if rc < 0
        msg = unsafe_string(ccall((:cfft_last_error, libcfft), Cstring, ()))
        rc == -1 ? throw(ArgumentError(msg)) : error("CatabraFFT: [$rc] $msg")
    end
    return rc
end

function context()
    lock(LOCK) do
        if CTX[] == C_NULL
            h = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:cfft_init, libcfft), Cint, (Cint, Ptr{Ptr{Cvoid}}), 0, h))
            CTX[] = h[]
        end
        CTX[]
    end
end

cfft_dtype(::Type{Float32}) = Cint(0)
cfft_dtype(::Type{Float64}) = Cint(1)
cfft_dtype(::Type{Float16}) = Cint(2)  # ComplexF16 storage, Float32 arithmetic (complex plans only)
cfft_dtype(::Type{T}) where {T} = throw(ArgumentError("CatabraFFT on MI355X supports ComplexF16, ComplexF32 and ComplexF64, got Complex{$T}"))

const PLANS = Dict{Tuple{Int,Int,DataType,FLAG},Ptr{Cvoid}}()

function plan_handle(n::Int, batch::Int, ::Type{T}, flag::FLAG) where {T<:AbstractFloat}
    lock(LOCK) do
        get!(PLANS, (n, batch, T, flag)) do
            h = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:cfft_plan_1d, libcfft), Cint,
                        (Ptr{Cvoid}, Int64, Int64, Int64, Int64, Cint, Cint, Ptr{Ptr{Cvoid}}),
                        context(), n, batch, n, n, cfft_dtype(T), Cint(flag), h))
            h[]
        end
    end
end

function exec_host!(y::Vector{Complex{T}}, x::Vector{Complex{T}}, dir::Cint, normalize::Bool,
                    flag::FLAG=NO_FLAG) where {T<:AbstractFloat}
    n = length(x)
    length(y) == n || throw(DimensionMismatch("output length $(length(y)) != $n"))
    n == 0 && return y
    h = plan_handle(n, 1, T, flag)
    GC.@preserve x y check(ccall((:cfft_exec_host, libcfft), Cint,
                                 (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint),
                                 h, pointer(x), pointer(y), dir, Cint(normalize)))
    return y
end

# Device-pointer execution (AMDGPU.jl): `stream` is the hipStream_t handle of the queue.
function exec_device!(y_ptr::Ptr{Cvoid}, x_ptr::Ptr{Cvoid}, n::Int, batch::Int, ::Type{T},
                      dir::Cint, normalize::Bool, stream::Ptr{Cvoid}=C_NULL;
                      flag::FLAG=NO_FLAG) where {T<:AbstractFloat}
    h = plan_handle(n, batch, T, flag)
    check(ccall((:cfft_exec, libcfft), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}),
                h, x_ptr, y_ptr, dir, Cint(normalize), stream))
end

_dense(x::AbstractVector{Complex{T}}) where {T} = x isa Vector{Complex{T}} ? x : Vector{Complex{T}}(x)

"""fft(x): forward DFT, a new vector (CatabraFFT.jl:65-72)."""
function fft(x::AbstractVector{Complex{T}})::AbstractVector{Complex{T}} where {T<:AbstractFloat}
    exec_host!(similar(_dense(x)), _dense(x), FORWARD, false)
end

"""ifft(x) = conj(fft(conj x)) / n (CatabraFFT.jl:88-101)."""
function ifft(x::AbstractVector{Complex{T}})::AbstractVector{Complex{T}} where {T<:AbstractFloat}
    exec_host!(similar(_dense(x)), _dense(x), BACKWARD, true)
end

"""bfft(x) = conj(fft(conj x)) (CatabraFFT.jl:103-115)."""
function bfft(x::AbstractVector{Complex{T}})::AbstractVector{Complex{T}} where {T<:AbstractFloat}
    exec_host!(similar(_dense(x)), _dense(x), BACKWARD, false)
end

"""Plan handle with the fields of spells.jl:18-28."""
struct Spell{T} <: AbstractFFTs.Plan{T}
    n::Int
    datatype::Type{T}
    flag::FLAG
    inverse::Union{Nothing,Spell{T}}
end
Spell(n::Int, ::Type{T}, flag::FLAG) where {T<:AbstractFloat} = Spell{T}(n, T, flag, nothing)

function plan_fft(x::AbstractVector{Complex{T}}, flags::FLAG)::Spell{T} where {T<:AbstractFloat}
    plan_handle(length(x), 1, T, flags)  # plans (and, for MEASURE / ENCHANT, tunes) now
    Spell(length(x), T, flags)
end
AbstractFFTs.plan_fft(x::AbstractVector{Complex{T}}, region=1:1; flags::FLAG=NO_FLAG) where {T<:AbstractFloat} =
    plan_fft(x, flags)

function mul!(y::AbstractVector{Complex{T}}, p::Spell, x::AbstractVector{Complex{T}}) where {T}
    length(x) == p.n || throw(DimensionMismatch("plan is for n=$(p.n), got $(length(x))"))
    yd = y isa Vector{Complex{T}} ? y : similar(_dense(x))
    exec_host!(yd, _dense(x), FORWARD, false, p.flag)
    yd === y || copyto!(y, yd)
    return y
end

Base.:*(p::Spell, x::AbstractVector{Complex{T}}) where {T} = mul!(similar(x), p, x)
(p::Spell{T})(x::AbstractVector{Complex{T}}) where {T} = fft(x)

# ---- real-input transforms (rfft / brfft / irfft): cfft_plan_1d_real; the reference's
# plan_rfft / plan_brfft stubs (CatabraFFT.jl:199-208) construct nothing usable.
const RPLANS = Dict{Tuple{Int,DataType},Ptr{Cvoid}}()

function real_plan(n::Int, ::Type{T}) where {T<:AbstractFloat}
    lock(LOCK) do
        get!(RPLANS, (n, T)) do
            h = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:cfft_plan_1d_real, libcfft), Cint,
                        (Ptr{Cvoid}, Int64, Int64, Int64, Int64, Cint, Cint, Ptr{Ptr{Cvoid}}),
                        context(), n, 1, n, n ÷ 2 + 1, cfft_dtype(T), Cint(NO_FLAG), h))
            h[]
        end
    end
end

"""rfft(x): the n÷2+1 non-negative-frequency terms of the DFT of real x."""
function rfft(x::AbstractVector{T}) where {T<:Union{Float32,Float64}}
    xd = x isa Vector{T} ? x : Vector{T}(x)
    y = Vector{Complex{T}}(undef, length(x) ÷ 2 + 1)
    GC.@preserve xd y check(ccall((:cfft_exec_host, libcfft), Cint,
                                  (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint),
                                  real_plan(length(x), T), pointer(xd), pointer(y), FORWARD, 0))
    return y
end

function _c2r(X::AbstractVector{Complex{T}}, d::Int, normalize::Bool) where {T<:Union{Float32,Float64}}
    length(X) == d ÷ 2 + 1 || throw(DimensionMismatch("$(length(X)) values do not match d=$d"))
    Xd = X isa Vector{Complex{T}} ? X : Vector{Complex{T}}(X)
    y = Vector{T}(undef, d)
    GC.@preserve Xd y check(ccall((:cfft_exec_host, libcfft), Cint,
                                  (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint),
                                  real_plan(d, T), pointer(Xd), pointer(y), BACKWARD, Cint(normalize)))
    return y
end
"""brfft(X, d): unnormalised inverse of rfft to d reals."""
brfft(X::AbstractVector{Complex{T}}, d::Integer) where {T} = _c2r(X, Int(d), false)
"""irfft(X, d): inverse of rfft to d reals."""
irfft(X::AbstractVector{Complex{T}}, d::Integer) where {T} = _c2r(X, Int(d), true)

"""empty_cache(): drops plans and staging buffers (CatabraFFT.jl:39-43)."""
function empty_cache()
    lock(LOCK) do
        for h in values(PLANS)
            ccall((:cfft_plan_destroy, libcfft), Cint, (Ptr{Cvoid},), h)
        end
        empty!(PLANS)
        for h in values(RPLANS)
            ccall((:cfft_plan_destroy, libcfft), Cint, (Ptr{Cvoid},), h)
        end
        empty!(RPLANS)
        CTX[] == C_NULL || check(ccall((:cfft_empty_cache, libcfft), Cint, (Ptr{Cvoid},), CTX[]))
    end
    nothing
end

end # module
