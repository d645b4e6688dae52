# SPECTreconAMDChainRulesExt.jl -- ChainRulesCore rules for the device
# projector (extension triggered by SPECTreconAMD, SPECTrecon and
# ChainRulesCore together).
#
# - project(x, plan): the pullback of a view cotangent is its back-projection,
#   as the reference's learned-reconstruction example differentiates the
#   projector in x (docs/lit/examples/6-dl.jl:152-156, a Zygote @adjoint).
# - project_mu(x, mu, plan): the projection with the attenuation map as an
#   argument; the mu part of the pullback is the library's vector-Jacobian
#   product through the exponential cumulative sum of the attenuation
#   (sr_project_vjp_mu, recomputing the rotated image instead of storing it).
module SPECTreconAMDChainRulesExt

using ChainRulesCore: ChainRulesCore, NoTangent, @thunk, unthunk
using SPECTrecon: SPECTrecon
using SPECTreconAMD: SPECTplanAMD, project_mu, project_vjp_mu, dev_backproject!

function ChainRulesCore.rrule(::typeof(SPECTrecon.project), image::AbstractArray{<:Real,3}, plan::SPECTplanAMD;
                              index::AbstractVector{<:Int} = 1:plan.nview)
    views = SPECTrecon.project(image, plan; index)
    function project_pullback(dviews)
        dx = @thunk dev_backproject!(zeros(Float32, plan.imgsize), Array{Float32}(unthunk(dviews)), plan, index)
        return NoTangent(), dx, NoTangent()
    end
    return views, project_pullback
end

function ChainRulesCore.rrule(::typeof(project_mu), image::AbstractArray{<:Real,3}, mumap::AbstractArray{<:Real,3},
                              plan::SPECTplanAMD; index::AbstractVector{<:Int} = 1:plan.nview)
    views = project_mu(image, mumap, plan; index)   # (the plan now holds mumap)
    function project_mu_pullback(dviews)
        d = Array{Float32}(unthunk(dviews))
        dx = @thunk dev_backproject!(zeros(Float32, plan.imgsize), d, plan, index)
        dmu = @thunk project_vjp_mu(image, d, plan; index)
        return NoTangent(), dx, dmu, NoTangent()
    end
    return views, project_mu_pullback
end

end # module
