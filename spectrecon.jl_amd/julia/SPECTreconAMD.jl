# SPECTreconAMD.jl -- Julia binding of libspectrecon_amd.so (include/spectrecon_amd.h)
#
# The drop-in for SPECTrecon.jl's projector path on an MI355X: the same
# names and argument meaning as SPECTrecon (SPECTplan, project!, backproject!,
# project, backproject, mlem!, osem!, Ablock), dispatching on a device plan.
# Julia arrays are column-major like the C-ABI; view indices are converted to
# 0-based Int32.  Julia is not installed where this repository is built and
# tested, so this module is exercised only through the same C-ABI from Python.
module SPECTreconAMD

using LinearAlgebra: mul!
using LinearMapsAA: LinearMapAA, LinearMapAO

export SPECTplanAMD, project!, backproject!, project, backproject, mlem!, osem!, Ablock, osem_strided!, gradient,
       autotune!, set_graph!, project_vjp_mu

const libspect = joinpath(@__DIR__, "..", "lib", "libspectrecon_amd.so")

const SR_OK, SR_ERR_ARG, SR_ERR_DIM = 0, 1, 2
const SR_MEM_HOST = Int32(0)

struct SRPlanDesc                      # sr_plan_desc
    nx::Int32; ny::Int32; nz::Int32
    px::Int32; pz::Int32; nview::Int32
    dy::Float32
    interpmeth::Int32; mode::Int32; conv::Int32
end

# status -> the reference's exception types and messages
function check(rc::Integer)
    rc == SR_OK && return nothing
    msg = unsafe_string(ccall((:sr_last_error, libspect), Cstring, ()))
    rc == SR_ERR_ARG && throw(msg)                     # throw(String) as src/spectplan.jl:76-93
    rc == SR_ERR_DIM && throw(DimensionMismatch(msg))
    error("libspectrecon_amd [$rc]: $msg")
end

mutable struct SPECTplanAMD
    handle::Ptr{Cvoid}
    imgsize::NTuple{3,Int}
    px::Int; pz::Int; nview::Int
    viewangle::Vector{Float32}
    interpmeth::Symbol; mode::Symbol
    dy::Float32
end

# SPECTplan(mumap, psfs, dy; T, viewangle, interpmeth, nthread, mode)  (src/spectplan.jl:58-67)
function SPECTplanAMD(mumap::Array{<:Real,3}, psfs::Array{<:Real,4}, dy::Real;
        T::Type{<:AbstractFloat} = Float32,
        viewangle::AbstractVector{<:Real} = (0:size(psfs, 4) - 1) / size(psfs, 4) * T(2π),
        interpmeth::Symbol = :two, nthread::Int = Threads.nthreads(), mode::Symbol = :fast)
    T === Float32 || throw("only T=Float32 runs on the device")
    (nthread == Threads.nthreads()) || throw("bad nthread")
    mumap .= T.(mumap); psfs .= T.(psfs)               # as src/spectplan.jl:71-72
    mu = Array{Float32}(mumap); ps = Array{Float32}(psfs)
    nx, ny, nz = size(mu); px, pz, _, nview = size(ps)
    im = interpmeth === :one ? 1 : interpmeth === :two ? 2 : 0
    md = mode === :fast ? 0 : mode === :mem ? 1 : -1
    desc = Ref(SRPlanDesc(nx, ny, nz, px, pz, nview, Float32(dy), im, md, 0))
    va = Vector{Float32}(viewangle)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sr_plan_create, libspect), Cint,
                (Ref{SRPlanDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{Ptr{Cvoid}}),
                desc, mu, ps, va, h))
    plan = SPECTplanAMD(h[], (nx, ny, nz), px, pz, nview, va, interpmeth, mode, Float32(dy))
    finalizer(p -> ccall((:sr_plan_destroy, libspect), Cint, (Ptr{Cvoid},), p.handle), plan)
    return plan
end

idx32(index) = Int32.(collect(index) .- 1)             # 1-based Julia views -> 0-based

# project!(views, image, plan; index)  (src/project.jl:143-163)
function project!(views::AbstractArray{Float32,3}, image::AbstractArray{Float32,3}, plan::SPECTplanAMD;
        index::AbstractVector{<:Int} = 1:plan.nview)
    ix = idx32(index)
    check(ccall((:sr_project, libspect), Cint, (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Int32, Int32),
                plan.handle, views, image, ix, length(ix), SR_MEM_HOST))
    return views
end

# backproject!(image, views, plan; index)  (src/backproject.jl:132-162)
function backproject!(image::AbstractArray{Float32,3}, views::AbstractArray{Float32,3}, plan::SPECTplanAMD;
        index::AbstractVector{<:Int} = 1:plan.nview)
    ix = idx32(index)
    check(ccall((:sr_backproject, libspect), Cint, (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Int32, Int32),
                plan.handle, image, views, ix, length(ix), SR_MEM_HOST))
    return image
end

project(image, plan::SPECTplanAMD; kwargs...) =
    project!(Array{Float32}(undef, plan.imgsize[1], plan.imgsize[3], plan.nview), image, plan; kwargs...)
backproject(views, plan::SPECTplanAMD; kwargs...) =
    backproject!(zeros(Float32, plan.imgsize), views, plan; kwargs...)

# The plan as the LinearMapAA the reference's own mlem!/osem! consume
# (docs/lit/examples/4-mlem.jl:78-82, src/ml-os-em.jl:70-84)
function Ablock(plan::SPECTplanAMD, nblocks::Int)
    nx, ny, nz = plan.imgsize
    rem(plan.nview, nblocks) == 0 || throw("nview must be divisible by nblocks!")
    map(1:nblocks) do nb
        viewidx = nb:nblocks:plan.nview
        odim = (nx, nz, length(viewidx))
        LinearMapAA((y, x) -> project!(y, x, plan; index = viewidx),
                    (x, y) -> backproject!(x, y, plan; index = viewidx),
                    (prod(odim), nx * ny * nz); T = Float32, odim, idim = (nx, ny, nz))
    end
end

# device-resident ML-EM / OS-EM (src/ml-os-em.jl:21-49, :97-131)
function mlem!(out::Array{Float32,3}, x0::Array{Float32,3}, ynoisy::Array{Float32,3},
        background::Array{Float32,3}, plan::SPECTplanAMD; niter::Int = 20)
    check(ccall((:sr_mlem, libspect), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Int32),
                plan.handle, out, x0, ynoisy, background, niter, SR_MEM_HOST))
    return out
end

function osem!(out::Array{Float32,3}, x0::Array{Float32,3}, ynoisy::Array{Float32,3},
        background::Array{Float32,3}, plan::SPECTplanAMD, nblocks::Int; niter::Int = 16)
    check(ccall((:sr_osem, libspect), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Int32, Int32),
                plan.handle, out, x0, ynoisy, background, nblocks, niter, SR_MEM_HOST))
    return out
end

# OS-EM over strided subsets that need not divide the views (no reference
# counterpart: the batched / served use)
function osem_strided!(out::Array{Float32,3}, x0::Array{Float32,3}, ynoisy::Array{Float32,3},
        background::Array{Float32,3}, plan::SPECTplanAMD, nblocks::Int; niter::Int = 16)
    check(ccall((:sr_osem_strided, libspect), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Int32, Int32),
                plan.handle, out, x0, ynoisy, background, nblocks, niter, SR_MEM_HOST))
    return out
end

# (L, dL/dx, dL/dmu) for L = ||A(x, mu) - y||^2 over the indexed views (no
# reference counterpart).  A ChainRulesCore rrule for a mu-dependent projector
# calls it with y = A x - dy/2 for the cotangent dy.
function gradient(image::Array{Float32,3}, y::Array{Float32,3}, plan::SPECTplanAMD;
                  index::AbstractVector{<:Integer} = 1:plan.nview)
    ix = Int32.(collect(index) .- 1)
    gx, gmu, loss = similar(image), similar(image), Ref{Float64}(0)
    check(ccall((:sr_grad, libspect), Cint,
                (Ptr{Cvoid}, Ref{Float64}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32},
                 Int32, Int32),
                plan.handle, loss, gx, gmu, image, y, ix, length(ix), SR_MEM_HOST))
    return loss[], gx, gmu
end

# d<dy, A(x, mu)>/dmu for view cotangents dy (the mu part of the projector's
# vector-Jacobian product; the image part is backproject(dy)).
function project_vjp_mu(image::Array{Float32,3}, dy::Array{Float32,3}, plan::SPECTplanAMD;
                        index::AbstractVector{<:Integer} = 1:plan.nview)
    ix = Int32.(collect(index) .- 1)
    gmu = similar(image)
    check(ccall((:sr_project_vjp_mu, libspect), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Int32, Int32),
                plan.handle, gmu, image, dy, ix, length(ix), SR_MEM_HOST))
    return gmu
end

# Kernel choices timed on this device for batches of nsub views (the OS-EM
# subset size); returns the JSON record of every candidate (no reference
# counterpart).  nsub <= 0 restores the built-in choices.
function autotune!(plan::SPECTplanAMD, nsub::Integer)
    buf = zeros(UInt8, 16384)
    check(ccall((:sr_plan_autotune, libspect), Cint, (Ptr{Cvoid}, Int32, Ptr{UInt8}, Csize_t),
                plan.handle, nsub, buf, length(buf)))
    return unsafe_string(pointer(buf))
end

# Replay mlem!/osem! kernel sequences from HIP graphs (no reference counterpart).
function set_graph!(plan::SPECTplanAMD, enable::Bool = true)
    check(ccall((:sr_plan_set_graph, libspect), Cint, (Ptr{Cvoid}, Int32), plan.handle, enable ? 1 : 0))
    return plan
end

end # module
