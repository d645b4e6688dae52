# SPECTreconAMDExt.jl -- the package extension that makes SPECTrecon.jl's own
# functions run on a device plan: methods of SPECTrecon.project!,
# backproject!, project, backproject, mlem!, osem! and Ablock for
# SPECTreconAMD.SPECTplanAMD, and of imrotate!, imrotate_adj!, fft_conv!,
# fft_conv_adj! and fft_conv_adj2! for PlanRotateAMD / PlanPSFAMD.
# In SPECTrecon's Project.toml:
#   [weakdeps]    SPECTreconAMD = "<uuid>"
#   [extensions]  SPECTreconAMDExt = "SPECTreconAMD"
# Callers keep SPECTrecon's names and signatures; operators that are not
# device plans fall through to the reference's own loops.
module SPECTreconAMDExt

using SPECTrecon: SPECTrecon, RealU
using SPECTreconAMD: SPECTreconAMD, SPECTplanAMD, AblockAMD, dev_project!, dev_backproject!, dev_em!, dev_em_chat!,
                     PlanRotateAMD, PlanPSFAMD, dev_imrotate!, dev_fft_conv!, dev_fft_conv3!
using LinearMapsAA: LinearMapAA, LinearMapAO

# ---- project! (src/project.jl) ----------------------------------------------------
# project!(views, image, plan; index)  (src/project.jl:143-163)
SPECTrecon.project!(views::AbstractArray{<:RealU,3}, image::AbstractArray{<:RealU,3}, plan::SPECTplanAMD;
                    index::AbstractVector{<:Int} = 1:plan.nview) = dev_project!(views, image, plan, index)

# project!(view, image, plan, viewidx)  (src/project.jl:11-69)
function SPECTrecon.project!(view::AbstractMatrix{<:RealU}, image::AbstractArray{<:RealU,3}, plan::SPECTplanAMD,
                             viewidx::Int)
    v = Array{Float32}(undef, size(view, 1), size(view, 2), 1)
    dev_project!(v, image, plan, viewidx:viewidx)
    copyto!(view, (@view v[:, :, 1]))
    return view
end

# project!(view, image, plan, buffer_id, viewidx)  (src/project.jl:77-135); the
# buffer_id picks a thread's work buffers there, the device plan owns its own
SPECTrecon.project!(view::AbstractMatrix{<:RealU}, image::AbstractArray{<:RealU,3}, plan::SPECTplanAMD,
                    buffer_id::Int, viewidx::Int) = SPECTrecon.project!(view, image, plan, viewidx)

# project(image, plan; index)  (src/project.jl:166-181)
function SPECTrecon.project(image::AbstractArray{<:RealU,3}, plan::SPECTplanAMD;
                            index::AbstractVector{<:Int} = 1:plan.nview)
    views = Array{Float32}(undef, plan.imgsize[1], plan.imgsize[3], length(index))
    return dev_project!(views, image, plan, index)
end

# ---- backproject! (src/backproject.jl) -------------------------------------------
# backproject!(image, views, plan; index)  (src/backproject.jl:132-162): image
# is overwritten, as the reference zeroes it
SPECTrecon.backproject!(image::AbstractArray{<:RealU,3}, views::AbstractArray{<:RealU,3}, plan::SPECTplanAMD;
                        index::AbstractVector{<:Int} = 1:plan.nview) = dev_backproject!(image, views, plan, index)

# backproject!(image, view, plan, viewidx)  (src/backproject.jl:9-62): writes image
SPECTrecon.backproject!(image::AbstractArray{<:RealU,3}, view::AbstractMatrix{<:RealU}, plan::SPECTplanAMD,
                        viewidx::Int) = dev_backproject!(image, reshape(Array{Float32}(view), size(view)..., 1), plan,
                                                         viewidx:viewidx)

# backproject!(image, view, plan, buffer_id, viewidx)  (src/backproject.jl:69-124):
# adds the view's back-projection to image (:121)
function SPECTrecon.backproject!(image::AbstractArray{<:RealU,3}, view::AbstractMatrix{<:RealU}, plan::SPECTplanAMD,
                                 buffer_id::Int, viewidx::Int)
    one = zeros(Float32, plan.imgsize)
    SPECTrecon.backproject!(one, view, plan, viewidx)
    image .+= one
    return image
end

# backproject(views, plan; index)  (src/backproject.jl:170-178)
SPECTrecon.backproject(views::AbstractArray{<:RealU,3}, plan::SPECTplanAMD;
                       index::AbstractVector{<:Int} = 1:plan.nview) =
    dev_backproject!(zeros(Float32, plan.imgsize), views, plan, index)

# ---- Ablock, mlem!, osem! (src/ml-os-em.jl) ----------------------------------------
# Ablock(plan, nblocks)  (src/ml-os-em.jl:70-84): the subset operators,
# remembering the plan
SPECTrecon.Ablock(plan::SPECTplanAMD, nblocks::Int) = AblockAMD(plan, nblocks)

# osem!(out, x0, ynoisy, background, Ab; niter, chat)  (src/ml-os-em.jl:97-131)
# on the device for Ablock(plan, nblocks) of a device plan, the same checks
# and messages (background > 0 first)
function SPECTrecon.osem!(out::AbstractArray, x0::AbstractArray, ynoisy::AbstractArray, background::AbstractArray,
                          Ab::AblockAMD; niter::Int = 16, chat::Bool = false)
    all(>(0), background) || throw("need background > 0")
    size(out) != size(x0) && throw(DimensionMismatch("size out and x0 not match"))
    chat && return dev_em_chat!(out, x0, ynoisy, background, Ab.plan, Ab.nblocks, niter)
    return dev_em!(out, x0, ynoisy, background, Ab.plan, Ab.nblocks, niter)
end

# mlem!(out, x0, ynoisy, background, A; niter, chat)  (src/ml-os-em.jl:21-49) on
# the device when A is the operator of all a device plan's views
# (SPECTreconAMD.operator(plan)); any other LinearMapAA runs the reference's loop
function SPECTrecon.mlem!(out::AbstractArray, x0::AbstractArray, ynoisy::AbstractArray, background::AbstractArray,
                          A::LinearMapAA; niter::Int = 20, chat::Bool = false)
    prop = A._prop
    if !(prop isa NamedTuple && haskey(prop, :spect_amd_plan) && prop.spect_amd_full)
        return invoke(SPECTrecon.mlem!,
                      Tuple{AbstractArray,AbstractArray,AbstractArray,AbstractArray,Union{AbstractArray,LinearMapAO}},
                      out, x0, ynoisy, background, A; niter, chat)
    end
    all(>(0), background) || throw("need background > 0")
    size(out) == size(x0) || throw(DimensionMismatch("size out and x0 not match"))
    size(ynoisy) == size(background) || throw(DimensionMismatch("size ynoisy and background"))
    chat && return dev_em_chat!(out, x0, ynoisy, background, prop.spect_amd_plan, 1, niter)
    return dev_em!(out, x0, ynoisy, background, prop.spect_amd_plan, 1, niter)
end

# (mlem(x0, ...) and osem(x0, ...), src/ml-os-em.jl:61, :143, call these mlem! / osem!)

# ---- unit operators for device plans ------------------------------------------------
# imrotate! / imrotate_adj! (src/rotatez.jl:240-416): one slice with a plan, a
# 3-D stack with a vector of plans (:472-558; the device takes every slice at once)
SPECTrecon.imrotate!(output::AbstractMatrix{<:RealU}, img::AbstractMatrix{<:RealU}, θ::RealU,
                     plan::PlanRotateAMD) = dev_imrotate!(output, img, θ, plan, false)
SPECTrecon.imrotate_adj!(output::AbstractMatrix{<:RealU}, img::AbstractMatrix{<:RealU}, θ::RealU,
                         plan::PlanRotateAMD) = dev_imrotate!(output, img, θ, plan, true)
SPECTrecon.imrotate!(output::AbstractArray{<:RealU,3}, image3::AbstractArray{<:RealU,3}, θ::RealU,
                     plans::Vector{PlanRotateAMD}) = dev_imrotate!(output, image3, θ, plans[1], false)
SPECTrecon.imrotate_adj!(output::AbstractArray{<:RealU,3}, image3::AbstractArray{<:RealU,3}, θ::RealU,
                         plans::Vector{PlanRotateAMD}) = dev_imrotate!(output, image3, θ, plans[1], true)

# fft_conv! / fft_conv_adj! (src/fft_convolve.jl:33-130) and the 3-D forms (:160-209)
SPECTrecon.fft_conv!(output::AbstractMatrix{<:RealU}, img::AbstractMatrix{<:RealU}, ker::AbstractMatrix{<:RealU},
                     plan::PlanPSFAMD) = dev_fft_conv!(output, img, ker, plan, false)
SPECTrecon.fft_conv_adj!(output::AbstractMatrix{<:RealU}, img::AbstractMatrix{<:RealU},
                         ker::AbstractMatrix{<:RealU}, plan::PlanPSFAMD) = dev_fft_conv!(output, img, ker, plan, true)
function SPECTrecon.fft_conv!(output::AbstractArray{<:RealU,3}, image3::AbstractArray{<:RealU,3},
                              ker3::AbstractArray{<:RealU,3}, plans::Vector{PlanPSFAMD})
    size(output) == size(image3) || throw(DimensionMismatch())
    return dev_fft_conv3!(output, image3, ker3, plans[1], 0)
end
function SPECTrecon.fft_conv_adj!(output::AbstractArray{<:RealU,3}, image3::AbstractArray{<:RealU,3},
                                  ker3::AbstractArray{<:RealU,3}, plans::Vector{PlanPSFAMD})
    size(output) == size(image3) || throw(DimensionMismatch())
    return dev_fft_conv3!(output, image3, ker3, plans[1], 1)
end

# fft_conv_adj2!(output, image2, ker3, plans)  (src/fft_convolve.jl:217-239):
# like the reference it writes planes iy = 1:size(image2, 2) of output (all of
# them when nz == ny; a BoundsError when nz > ny)
function SPECTrecon.fft_conv_adj2!(output::AbstractArray{<:RealU,3}, image2::AbstractMatrix{<:RealU},
                                   ker3::AbstractArray{<:RealU,3}, plans::Vector{PlanPSFAMD})
    size(output, 1) == size(image2, 1) || throw("size 1")
    size(output, 3) == size(image2, 2) || throw("size 2")
    nz = size(image2, 2)
    nz <= size(output, 2) || throw(BoundsError(output, (1, nz, 1)))
    nz == size(output, 2) && return dev_fft_conv3!(output, image2, ker3, plans[1], 2)
    part = Array{Float32}(undef, size(output, 1), nz, nz)
    dev_fft_conv3!(part, image2, ker3, plans[1], 2)
    output[:, 1:nz, :] .= part
    return output
end

end # module
