# SPECTreconAMD.jl -- Julia binding of libspectrecon_amd.so (include/spectrecon_amd.h)
#
# The ccall layer of the MI355X drop-in for SPECTrecon.jl's projector path.
# This module defines the device plan `SPECTplanAMD` (the counterpart of
# SPECTplan, src/spectplan.jl:35-138), the `AblockAMD` subset vector, and
# plan-level tools; it exports no name that SPECTrecon exports.  The
# reference's own functions (project!, backproject!, project, backproject,
# mlem!, osem!, Ablock) get their methods for device plans from the package
# extension ext/SPECTreconAMDExt.jl, so user code keeps calling SPECTrecon's
# names; ext/SPECTreconAMDChainRulesExt.jl adds the ChainRulesCore rrules.
#
# Julia arrays are column-major like the C-ABI; view indices are converted to
# 0-based Int32.  Julia is not installed where this repository is built and
# tested, so this module is exercised only through the same C-ABI from Python.
module SPECTreconAMD

using LinearMapsAA: LinearMapAA, LinearMapAO

export SPECTplanAMD, AblockAMD, operator, osem_strided!, gradient, project_vjp_mu, project_mu, set_mumap!,
       autotune!, set_graph!, PlanRotateAMD, PlanPSFAMD

const libspect = joinpath(@__DIR__, "..", "lib", "libspectrecon_amd.so")

const SR_OK, SR_ERR_ARG, SR_ERR_DIM = 0, 1, 2
const SR_MEM_HOST = Int32(0)
const SR_CONV = Dict(:auto => 0, :direct => 1, :separable => 2, :fft => 3)

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
    T::Type{Float32}
end

# SPECTplan(mumap, psfs, dy; T, viewangle, interpmeth, nthread, mode)  (src/spectplan.jl:58-67);
# `conv` (:auto, :direct, :separable, :fft) picks the PSF method (no reference counterpart)
function SPECTplanAMD(mumap::Array{<:Real,3}, psfs::Array{<:Real,4}, dy::Real;
        T::Type{<:AbstractFloat} = Float32,
        viewangle::AbstractVector{<:Real} = (0:size(psfs, 4) - 1) / size(psfs, 4) * T(2π),
        interpmeth::Symbol = :two, nthread::Int = Threads.nthreads(), mode::Symbol = :fast,
        conv::Symbol = :auto)
    T === Float32 || throw("only T=Float32 runs on the device")
    (nthread == Threads.nthreads()) || throw("bad nthread")
    mumap .= T.(mumap); psfs .= T.(psfs)               # as src/spectplan.jl:71-72
    mu = Array{Float32}(mumap); ps = Array{Float32}(psfs)
    nx, ny, nz = size(mu); px, pz, _, nview = size(ps)
    length(viewangle) == nview || throw(DimensionMismatch("viewangle length must equal size(psfs, 4)"))
    im = interpmeth === :one ? 1 : interpmeth === :two ? 2 : 0
    md = mode === :fast ? 0 : mode === :mem ? 1 : -1
    desc = Ref(SRPlanDesc(nx, ny, nz, px, pz, nview, Float32(dy), im, md, get(SR_CONV, conv, -1)))
    va = Vector{Float32}(viewangle)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sr_plan_create, libspect), Cint,
                (Ref{SRPlanDesc}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ref{Ptr{Cvoid}}),
                desc, mu, ps, va, h))
    plan = SPECTplanAMD(h[], (nx, ny, nz), px, pz, nview, va, interpmeth, mode, Float32(dy), Float32)
    finalizer(p -> ccall((:sr_plan_destroy, libspect), Cint, (Ptr{Cvoid},), p.handle), plan)
    return plan
end

Base.show(io::IO, p::SPECTplanAMD) =
    print(io, "SPECTplanAMD{Float32} imgsize=$(p.imgsize) px=$(p.px) pz=$(p.pz) nview=$(p.nview) ",
          "interpmeth=:$(p.interpmeth) mode=:$(p.mode)")

idx32(index) = Int32.(collect(index) .- 1)             # 1-based Julia views -> 0-based

# Runs f on Float32 column-major copies of A when A is not one already, and
# copies the result back (views of larger arrays, other element types).
function dense(f, A::AbstractArray; out::Bool)
    A isa Array{Float32} && return f(A)
    B = Array{Float32}(A)
    r = f(B)
    out && copyto!(A, B)
    return r
end

# ---- device calls, sizes checked before any ccall -------------------------------
function dev_project!(views::AbstractArray{<:Real,3}, image::AbstractArray{<:Real,3}, plan::SPECTplanAMD,
                      index::AbstractVector{<:Integer})
    nx, ny, nz = plan.imgsize
    size(image) == plan.imgsize || throw(DimensionMismatch("image size"))
    size(views) == (nx, nz, length(index)) || throw(DimensionMismatch("views size"))
    ix = idx32(index)
    dense(views; out = true) do v
        dense(image; out = false) do x
            check(ccall((:sr_project, libspect), Cint,
                        (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Int32, Int32),
                        plan.handle, v, x, ix, length(ix), SR_MEM_HOST))
        end
    end
    return views
end

function dev_backproject!(image::AbstractArray{<:Real,3}, views::AbstractArray{<:Real,3}, plan::SPECTplanAMD,
                          index::AbstractVector{<:Integer})
    nx, ny, nz = plan.imgsize
    size(image) == plan.imgsize || throw(DimensionMismatch("image size"))
    size(views) == (nx, nz, length(index)) || throw(DimensionMismatch("views size"))
    ix = idx32(index)
    dense(image; out = true) do x
        dense(views; out = false) do v
            check(ccall((:sr_backproject, libspect), Cint,
                        (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Int32, Int32),
                        plan.handle, x, v, ix, length(ix), SR_MEM_HOST))
        end
    end
    return image
end

# ---- unit operators ----------------------------------------------------------------
# Device counterparts of PlanRotate (plan_rotate(nx; method), src/plan-rotate.jl:81-103)
# and PlanPSF (plan_psf(; nx, nz, px, pz), src/plan-psf.jl:88-109): the library
# keeps the buffers and tables, so a plan holds only what the checks need.
struct PlanRotateAMD
    nx::Int
    method::Symbol
end
PlanRotateAMD(nx::Int; method::Symbol = :two) =
    method in (:one, :two) ? PlanRotateAMD(nx, method) : throw("unknown rotation method $method")

struct PlanPSFAMD
    nx::Int
    nz::Int
    px::Int
    pz::Int
end
PlanPSFAMD(; nx::Int = 128, nz::Int = nx, px::Int = 1, pz::Int = px) = PlanPSFAMD(nx, nz, px, pz)

# imrotate! / imrotate_adj! of every [:, :, k] slice (src/rotatez.jl:240-416, :472-558)
function dev_imrotate!(output::AbstractArray{<:Real}, img::AbstractArray{<:Real}, θ::Real, plan::PlanRotateAMD,
                       adj::Bool)
    size(output) == size(img) || throw(DimensionMismatch())
    n = size(img, 1)
    (size(img, 2) == n && n == plan.nx) || throw(DimensionMismatch("square slices of the plan's size"))
    nz = ndims(img) == 3 ? size(img, 3) : 1
    interp = Int32(plan.method === :one ? 1 : 2)
    dense(output; out = true) do o
        dense(img; out = false) do x
            check(ccall((:sr_imrotate, libspect), Cint,
                        (Ptr{Float32}, Ptr{Float32}, Int32, Int32, Float32, Int32, Int32, Int32),
                        o, x, n, nz, Float32(θ), interp, Int32(adj), SR_MEM_HOST))
        end
    end
    return output
end

# fft_conv! / fft_conv_adj! of one plane (src/fft_convolve.jl:33-130) with its
# checks and messages
function dev_fft_conv!(output::AbstractMatrix{<:Real}, img::AbstractMatrix{<:Real}, ker::AbstractMatrix{<:Real},
                       plan::PlanPSFAMD, adj::Bool)
    size(output) == size(img) || throw("size output")
    size(img) == (plan.nx, plan.nz) || throw("size img")
    size(ker) == (plan.px, plan.pz) || throw("size ker $(size(ker)) $(plan.px) $(plan.pz)")
    dense(output; out = true) do o
        dense(img; out = false) do x
            dense(ker; out = false) do k
                check(ccall((:sr_fft_conv, libspect), Cint,
                            (Ptr{Float32}, Ptr{Float32}, Int32, Int32, Ptr{Float32}, Int32, Int32, Int32, Int32),
                            o, x, plan.nx, plan.nz, k, plan.px, plan.pz, Int32(adj), SR_MEM_HOST))
            end
        end
    end
    return output
end

# the 3-D forms (src/fft_convolve.jl:160-239), every plane in one call:
# form 0 fft_conv!, 1 fft_conv_adj! (image is [nx, ny, nz]), 2 fft_conv_adj2!
# (image is one [nx, nz] plane; output planes iy = 1:size(output, 2))
function dev_fft_conv3!(output::AbstractArray{<:Real,3}, image::AbstractArray{<:Real}, ker3::AbstractArray{<:Real,3},
                        plan::PlanPSFAMD, form::Int)
    nx, ny, nz = size(output)
    (nx, nz) == (plan.nx, plan.nz) || throw("size img")
    (size(ker3, 1), size(ker3, 2)) == (plan.px, plan.pz) || throw("size ker $(size(ker3)[1:2]) $(plan.px) $(plan.pz)")
    size(ker3, 3) >= ny || throw(BoundsError(ker3, (1, 1, ny)))
    k3 = size(ker3, 3) == ny ? ker3 : ker3[:, :, 1:ny]
    dense(output; out = true) do o
        dense(image; out = false) do x
            dense(k3; out = false) do k
                check(ccall((:sr_fft_conv3, libspect), Cint,
                            (Ptr{Float32}, Ptr{Float32}, Int32, Int32, Int32, Ptr{Float32}, Int32, Int32, Int32,
                             Int32),
                            o, x, nx, ny, nz, k, plan.px, plan.pz, Int32(form), SR_MEM_HOST))
            end
        end
    end
    return output
end

# x0/out: the plan's image size; y/b: [nx, nz, nview] (the library reads and
# writes those many floats)
function check_em(plan::SPECTplanAMD, out, x0, ynoisy, background)
    nx, ny, nz = plan.imgsize
    size(out) == plan.imgsize || throw(DimensionMismatch("size out and plan image size not match"))
    size(x0) == plan.imgsize || throw(DimensionMismatch("size x0 and plan image size not match"))
    size(ynoisy) == (nx, nz, plan.nview) || throw(DimensionMismatch("size ynoisy and plan views not match"))
    size(background) == size(ynoisy) || throw(DimensionMismatch("size ynoisy and background"))
    return nothing
end

# device-resident ML-EM / OS-EM over strided blocks nb:nblocks:nview
# (src/ml-os-em.jl:21-49 with nblocks = 1, :97-131 otherwise); `strided`
# allows nblocks that do not divide nview (sr_osem_strided)
function dev_em!(out, x0, ynoisy, background, plan::SPECTplanAMD, nblocks::Int, niter::Int; strided::Bool = false)
    all(>(0), background) || throw("need background > 0")
    check_em(plan, out, x0, ynoisy, background)
    dense(out; out = true) do o
        dense(x0; out = false) do x
            dense(ynoisy; out = false) do y
                dense(background; out = false) do b
                    if strided
                        check(ccall((:sr_osem_strided, libspect), Cint,
                                    (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Int32,
                                     Int32), plan.handle, o, x, y, b, nblocks, niter, SR_MEM_HOST))
                    elseif nblocks == 1
                        check(ccall((:sr_mlem, libspect), Cint,
                                    (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Int32),
                                    plan.handle, o, x, y, b, niter, SR_MEM_HOST))
                    else
                        check(ccall((:sr_osem, libspect), Cint,
                                    (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Int32, Int32,
                                     Int32), plan.handle, o, x, y, b, nblocks, niter, SR_MEM_HOST))
                    end
                end
            end
        end
    end
    return out
end

# EM with the reference's `chat` output: one device iteration per call (a
# call keeps no state besides the image, so the iterates are the same),
# showing (iter, extrema, time) before each as src/ml-os-em.jl:42 / :122 do
# (OS-EM: once per iteration, at its first block)
function dev_em_chat!(out, x0, ynoisy, background, plan::SPECTplanAMD, nblocks::Int, niter::Int)
    copyto!(out, x0)
    cur = Array{Float32}(x0)
    time0 = time()
    for iter in 1:niter
        if nblocks == 1
            @show iter, extrema(cur), time() - time0
        else
            nb = 1
            @show iter, nb, extrema(cur), time() - time0
        end
        dev_em!(out, cur, ynoisy, background, plan, nblocks, 1)
        copyto!(cur, out)
    end
    return out
end

# ---- operators ---------------------------------------------------------------
# The plan (or its views `index`) as the LinearMapAA the reference's loops
# consume (docs/lit/examples/4-mlem.jl:78-82), tagged with the plan so that
# SPECTrecon's mlem! runs the device-resident loop for it.
function operator(plan::SPECTplanAMD; index::AbstractVector{<:Int} = 1:plan.nview)
    nx, ny, nz = plan.imgsize
    odim = (nx, nz, length(index))
    full = collect(index) == collect(1:plan.nview)
    LinearMapAA((y, x) -> dev_project!(y, x, plan, index), (x, y) -> dev_backproject!(x, y, plan, index),
                (prod(odim), nx * ny * nz); T = Float32, odim, idim = (nx, ny, nz),
                prop = (spect_amd_plan = plan, spect_amd_full = full))
end

# Ablock(plan, nblocks) for a device plan (src/ml-os-em.jl:70-84): the subset
# operators nb:nblocks:nview, remembering the plan so that SPECTrecon's osem!
# runs the device-resident loop for them.
struct AblockAMD <: AbstractVector{LinearMapAO}
    plan::SPECTplanAMD
    nblocks::Int
    blocks::Vector{LinearMapAO}
end

function AblockAMD(plan::SPECTplanAMD, nblocks::Int)
    rem(plan.nview, nblocks) == 0 || throw("nview must be divisible by nblocks!")
    blocks = LinearMapAO[operator(plan; index = nb:nblocks:plan.nview) for nb in 1:nblocks]
    return AblockAMD(plan, nblocks, blocks)
end
Base.size(A::AblockAMD) = (A.nblocks,)
Base.getindex(A::AblockAMD, i::Int) = A.blocks[i]

# OS-EM over strided subsets that need not divide the views (no reference
# counterpart: the batched / served use)
osem_strided!(out::AbstractArray{<:Real,3}, x0::AbstractArray{<:Real,3}, ynoisy::AbstractArray{<:Real,3},
              background::AbstractArray{<:Real,3}, plan::SPECTplanAMD, nblocks::Int; niter::Int = 16) =
    dev_em!(out, x0, ynoisy, background, plan, nblocks, niter; strided = true)

# ---- mu: a new attenuation map, gradients -----------------------------------------
# The plan's mu-map replaced; every view's attenuation factors rebuilt on the device
function set_mumap!(plan::SPECTplanAMD, mumap::AbstractArray{<:Real,3})
    size(mumap) == plan.imgsize || throw(DimensionMismatch("mumap size"))
    mu = Array{Float32}(mumap)
    check(ccall((:sr_plan_set_mumap, libspect), Cint, (Ptr{Cvoid}, Ptr{Float32}, Int32), plan.handle, mu, SR_MEM_HOST))
    return plan
end

# A(x, mu): the projection with the plan's mu-map replaced by `mumap`; the
# function ChainRulesCore differentiates in both x and mu
function project_mu(image::AbstractArray{<:Real,3}, mumap::AbstractArray{<:Real,3}, plan::SPECTplanAMD;
                    index::AbstractVector{<:Int} = 1:plan.nview)
    set_mumap!(plan, mumap)
    views = Array{Float32}(undef, plan.imgsize[1], plan.imgsize[3], length(index))
    return dev_project!(views, image, plan, index)
end

# (L, dL/dx, dL/dmu) for L = ||A(x, mu) - y||^2 over the indexed views (no
# reference counterpart), y holding the indexed views in order
function gradient(image::AbstractArray{<:Real,3}, y::AbstractArray{<:Real,3}, plan::SPECTplanAMD;
                  index::AbstractVector{<:Integer} = 1:plan.nview)
    nx, ny, nz = plan.imgsize
    size(image) == plan.imgsize || throw(DimensionMismatch("image size"))
    size(y) == (nx, nz, length(index)) || throw(DimensionMismatch("views size"))
    ix = idx32(index)
    x, yy = Array{Float32}(image), Array{Float32}(y)
    gx, gmu, loss = similar(x), similar(x), Ref{Float64}(0)
    check(ccall((:sr_grad, libspect), Cint,
                (Ptr{Cvoid}, Ref{Float64}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32},
                 Int32, Int32),
                plan.handle, loss, gx, gmu, x, yy, ix, length(ix), SR_MEM_HOST))
    return loss[], gx, gmu
end

# d<dy, A(x, mu)>/dmu for view cotangents dy (the mu part of the projector's
# vector-Jacobian product; the image part is the back-projection of dy)
function project_vjp_mu(image::AbstractArray{<:Real,3}, dy::AbstractArray{<:Real,3}, plan::SPECTplanAMD;
                        index::AbstractVector{<:Integer} = 1:plan.nview)
    nx, ny, nz = plan.imgsize
    size(image) == plan.imgsize || throw(DimensionMismatch("image size"))
    size(dy) == (nx, nz, length(index)) || throw(DimensionMismatch("views size"))
    ix = idx32(index)
    x, d = Array{Float32}(image), Array{Float32}(dy)
    gmu = similar(x)
    check(ccall((:sr_project_vjp_mu, libspect), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Int32, Int32),
                plan.handle, gmu, x, d, ix, length(ix), SR_MEM_HOST))
    return gmu
end

# ---- tuning ----------------------------------------------------------------------
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
