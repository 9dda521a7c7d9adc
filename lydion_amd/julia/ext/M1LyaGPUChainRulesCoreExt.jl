"""
The ChainRules `rrule` of M1LyaGPU.run_MF, loaded by Julia's package-extension mechanism when
ChainRulesCore is in the environment (the reference's Project.toml does not list it).  The
pullback replays the same steps through m1lya_adjoint_run (reverse-mode gradient with √N
checkpointing) and returns the cotangents of the LyαArrays fields the step reads.
"""
module M1LyaGPUChainRulesCoreExt

using ChainRulesCore
using M1LyaGPU: M1LyaGPU, run_MF, adjoint_run!, ctxof, upload!, F_J, F_H_R, F_H_Z

function ChainRulesCore.rrule(::typeof(run_MF), Lyα, nsteps::Integer, c̃::Float64, Δt::Float64, Z_center::Float64;
                              BC_Z_bot::Symbol=:Reflect)
    # the state before the loop, re-uploaded so that the pullback replays the same steps
    J0, Hr0, HZ0 = copy(Lyα.J_Lyα), copy(Lyα.H_r_Lyα), copy(Lyα.H_Z_Lyα)
    MF = run_MF(Lyα, nsteps, c̃, Δt, Z_center; BC_Z_bot=BC_Z_bot)
    function run_MF_pullback(ȳ)
        ctx = ctxof(Lyα)
        for (f, A) in ((F_J, J0), (F_H_R, Hr0), (F_H_Z, HZ0))
            upload!(ctx, f, A)
        end
        seeds = vcat(reshape(collect(Float64, unthunk(ȳ)), 1, :), zeros(2, nsteps))
        _, _, gs, gc, gb = adjoint_run!(Lyα, nsteps, c̃, Δt, Z_center; BC_Z_bot=BC_Z_bot, mf_seed=seeds)
        dLyα = Tangent{typeof(Lyα)}(; J_Lyα=gs[:, :, :, 1], H_r_Lyα=gs[:, :, :, 2], H_Z_Lyα=gs[:, :, :, 3],
                                   α_Lyα=gc[:, :, 1], αc_Lyα=gc[:, :, 2], ∫dν_js_Lyα=gc[:, :, 3], ΔνD=gc[:, :, 4],
                                   α_F_Lyα=gb[:, :, :, 1], Hvoigt=gb[:, :, :, 2], ϕ_ν=gb[:, :, :, 3])
        return NoTangent(), dLyα, NoTangent(), NoTangent(), NoTangent(), NoTangent()
    end
    return MF, run_MF_pullback
end

end # module
