"""
    M1LyaGPU

Drop-in GPU backend for LYDION's Lyα M1 step.  Defines methods with the names
and signatures of the reference functions (src/UpdateLyaJandH.jl,
src/ConstantsAndFunctions.jl, src/FreqFactors.jl) whose bodies `ccall` into
libm1lya.so (include/m1lya.h).  The state lives on the device, keyed by the
LyαArrays object; `download_state!` copies it back at save/checkpoint points.

Usage in src/StaticCloudTest.jl (after GenerateLyαArrays / UpdateFreqFactors!):

    using .M1LyaGPU
    ctx = M1LyaGPU.attach!(LyaState, geometry, T, u_r_star, u_Z_star,
                           ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z, u_r_over_r)
    # the time loop body is unchanged: the methods below shadow the CPU ones
"""
module M1LyaGPU

export attach!, download_state!, upload_state!, edge_spectra, AdvectLyα!, LyαEmissionAbsorptionUpdate!,
       FrequencyUpdateJLyα_AllCells!, UpdateFreqFactors!, Update_f_Lyα_MF_L_Lyα!,
       Compute_f_esc_Lyα_Cylindrical, Reset_Δρu!, Update_e_Lyα!, BatchM1Lya, run_batch!, adjoint_run!,
       set_voigt_lut!, prepare_subcycling!, set_adjoint_mode!, run_until_steady!, GetRHDCFLTimeStep,
       FindRSL_LyαRT, absorption_substep!, run_MF, prepare_subcycling_chem!, run_batch_until_steady!

const libm1lya = joinpath(@__DIR__, "..", "..", "libm1lya.so")

# enum m1lya_field (include/m1lya.h)
const F_J, F_J_OLD, F_H_R, F_H_Z, F_H_R_OLD, F_H_Z_OLD = 0, 1, 2, 3, 4, 5
const F_ALPHA_F, F_HVOIGT, F_PHI_NU, F_FIJ_KP, F_FIJ_K, F_FIJ_KM = 6, 7, 8, 9, 10, 11
const F_ALPHA_LYA, F_ALPHA_C, F_PD, F_JS, F_T, F_DNU_D = 12, 13, 14, 15, 16, 17
const F_DUR_DR, F_DUR_DZ, F_DUZ_DR, F_DUZ_DZ, F_UR_OVER_R = 18, 19, 20, 21, 22
const F_DRHOU_R, F_DRHOU_Z, F_F_LYA_R, F_F_LYA_Z, F_E_LYA = 23, 24, 25, 26, 27
const F_UR_STAR, F_UZ_STAR = 28, 29

bc(s::Symbol) = s === :Reflect ? Cint(0) : s === :Outflow ? Cint(1) : error("unknown BC $s")

struct CGeometry           # m1lya_geometry
    r::Ptr{Float64}; Z::Ptr{Float64}; rCoM::Ptr{Float64}; ZCoM::Ptr{Float64}
    dV::Ptr{Float64}; inv_dV::Ptr{Float64}; A_r::Ptr{Float64}; A_Z::Ptr{Float64}; inv_r_cell::Ptr{Float64}
    dr_L::Ptr{Float64}; dr_R::Ptr{Float64}; inv_dr_i_im1::Ptr{Float64}; inv_dr_ip1_i::Ptr{Float64}
    inv_dr_ip2_ip1::Ptr{Float64}; dr_inner::NTuple{5,Float64}; dr_outer::NTuple{5,Float64}
    dZ_L::Ptr{Float64}; dZ_R::Ptr{Float64}; inv_dZ_j_jm1::Ptr{Float64}; inv_dZ_jp1_j::Ptr{Float64}
    inv_dZ_jp2_jp1::Ptr{Float64}; dZ_inner::NTuple{5,Float64}; dZ_outer::NTuple{5,Float64}
end

struct CFreqGrid           # m1lya_freq_grid
    nu::Ptr{Float64}; nuC::Ptr{Float64}; dnu::Ptr{Float64}; inv_dnu::Ptr{Float64}; inv_nuC::Ptr{Float64}
    kcore_min::Cint; kcore_max::Cint; recoil::Cint; flux_method::Cint; alpha_floor::Float64
end

const CONTEXTS = IdDict{Any,Ptr{Cvoid}}()

function check(rc::Cint, ctx::Ptr{Cvoid})
    rc == 0 && return nothing
    buf = Vector{UInt8}(undef, 1024)
    ccall((:m1lya_last_error, libm1lya), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Csize_t), ctx, buf, length(buf))
    error("libm1lya error $rc: " * unsafe_string(pointer(buf)))
end

ctxof(Lyα) = get(CONTEXTS, Lyα) do
    error("M1LyaGPU.attach! has not been called for this LyαArrays")
end

upload!(ctx, field, A::AbstractArray{Float64}) = GC.@preserve A check(
    ccall((:m1lya_upload, libm1lya), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Csize_t),
          ctx, field, A, length(A)), ctx)
download!(ctx, field, A::AbstractArray{Float64}) = GC.@preserve A check(
    ccall((:m1lya_download, libm1lya), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}, Csize_t),
          ctx, field, A, length(A)), ctx)
optr(A) = A === nothing ? Ptr{Float64}(C_NULL) : pointer(A)

# The pointers in these structs are only valid while the caller GC-preserves g / Lyα.
cgeometry(g) = CGeometry(pointer(g.r), pointer(g.Z), pointer(g.rCoM), pointer(g.ZCoM), pointer(g.ΔV),
                         pointer(g.inv_ΔV), pointer(g.A_r), pointer(g.A_Z), pointer(g.inv_r_cell),
                         pointer(g.Δr_L), pointer(g.Δr_R), pointer(g.inv_dr_i_im1), pointer(g.inv_dr_ip1_i),
                         pointer(g.inv_dr_ip2_ip1),
                         (g.Δr_L_InnerBoundary, g.Δr_R_InnerBoundary, g.inv_dr_i_im1_InnerBoundary,
                          g.inv_dr_ip1_i_InnerBoundary, g.inv_dr_ip2_ip1_InnerBoundary),
                         (g.Δr_L_OuterBoundary, g.Δr_R_OuterBoundary, g.inv_dr_i_im1_OuterBoundary,
                          g.inv_dr_ip1_i_OuterBoundary, g.inv_dr_ip2_ip1_OuterBoundary),
                         pointer(g.ΔZ_L), pointer(g.ΔZ_R), pointer(g.inv_dZ_j_jm1), pointer(g.inv_dZ_jp1_j),
                         pointer(g.inv_dZ_jp2_jp1),
                         (g.ΔZ_L_InnerBoundary, g.ΔZ_R_InnerBoundary, g.inv_dZ_j_jm1_InnerBoundary,
                          g.inv_dZ_jp1_j_InnerBoundary, g.inv_dZ_jp2_jp1_InnerBoundary),
                         (g.ΔZ_L_OuterBoundary, g.ΔZ_R_OuterBoundary, g.inv_dZ_j_jm1_OuterBoundary,
                          g.inv_dZ_jp1_j_OuterBoundary, g.inv_dZ_jp2_jp1_OuterBoundary))

cfreqgrid(Lyα) = CFreqGrid(pointer(Lyα.ν), pointer(Lyα.νC), pointer(Lyα.Δν), pointer(Lyα.inv_Δν),
                           pointer(Lyα.inv_νC), Lyα.kcore_min - 1, Lyα.kcore_max - 1, Lyα.Recoil,
                           Lyα.FluxMethod isa Val{:HLL} ? Cint(1) : Cint(0), Lyα.α_Lyα_floor)

"""Creates the device context of a LyαArrays + GeometryCylindrical and uploads
the static coefficients and the state (zero-copy: Julia arrays are [k,i,j])."""
function attach!(Lyα, g, T, u_r_star, u_Z_star, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z, u_r_over_r)
    N_ν, N_r, N_Z = size(Lyα.J_Lyα)
    ref = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:m1lya_create, libm1lya), Cint, (Cint, Cint, Cint, Cint, Ref{Ptr{Cvoid}}),
                N_ν, N_r, N_Z, 0, ref), C_NULL)
    ctx = ref[]
    GC.@preserve g begin
        cg = cgeometry(g)
        check(ccall((:m1lya_set_geometry, libm1lya), Cint, (Ptr{Cvoid}, Ref{CGeometry}), ctx, cg), ctx)
    end
    GC.@preserve Lyα begin
        fg = cfreqgrid(Lyα)
        check(ccall((:m1lya_set_freq_grid, libm1lya), Cint, (Ptr{Cvoid}, Ref{CFreqGrid}), ctx, fg), ctx)
    end
    for (f, A) in ((F_J, Lyα.J_Lyα), (F_J_OLD, Lyα.J_Lyα_old), (F_H_R, Lyα.H_r_Lyα), (F_H_Z, Lyα.H_Z_Lyα),
                   (F_H_R_OLD, Lyα.H_r_Lyα_old), (F_H_Z_OLD, Lyα.H_Z_Lyα_old), (F_ALPHA_F, Lyα.α_F_Lyα),
                   (F_HVOIGT, Lyα.Hvoigt), (F_PHI_NU, Lyα.ϕ_ν), (F_ALPHA_LYA, Lyα.α_Lyα), (F_ALPHA_C, Lyα.αc_Lyα),
                   (F_PD, Lyα.pd), (F_JS, Lyα.∫dν_js_Lyα), (F_DNU_D, Lyα.ΔνD), (F_T, T),
                   (F_DRHOU_R, Lyα.Δρu_Lyα_r), (F_DRHOU_Z, Lyα.Δρu_Lyα_Z), (F_UR_STAR, u_r_star),
                   (F_UZ_STAR, u_Z_star), (F_DUR_DR, ∂u_r_∂r), (F_DUR_DZ, ∂u_r_∂Z), (F_DUZ_DR, ∂u_Z_∂r),
                   (F_DUZ_DZ, ∂u_Z_∂Z), (F_UR_OVER_R, u_r_over_r))
        upload!(ctx, f, A)
    end
    CONTEXTS[Lyα] = ctx
    return ctx
end

"""Pushes the LyαArrays' J and H to the device, with the *_old arrays equal to them, and
recomputes e_Lyα there: the restart of src/RHDSolver_SingleStar.jl:1359-1372 after the reference
has read /Checkpoint's J_Lya, H_r_Lya, H_Z_Lya into LyαState."""
function upload_state!(Lyα)
    ctx = ctxof(Lyα)
    for (f, A) in ((F_J, Lyα.J_Lyα), (F_J_OLD, Lyα.J_Lyα), (F_H_R, Lyα.H_r_Lyα), (F_H_Z, Lyα.H_Z_Lyα),
                   (F_H_R_OLD, Lyα.H_r_Lyα), (F_H_Z_OLD, Lyα.H_Z_Lyα))
        upload!(ctx, f, A)
    end
    check(ccall((:m1lya_update_e_lya, libm1lya), Cint, (Ptr{Cvoid},), ctx), ctx)
    download!(ctx, F_E_LYA, Lyα.e_Lyα)
    return Lyα
end

"""The edge spectra SaveTimeStepHDF5 stores (src/SaveHDF5Data_StaticCloudTests.jl:392-420), from the
device state: (raw, normalised), each [N_edge, N_ν]."""
function edge_spectra(Lyα)
    ctx = ctxof(Lyα)
    ne = ccall((:m1lya_edge_count, libm1lya), Cint, (Ptr{Cvoid},), ctx)
    N_ν = size(Lyα.J_Lyα, 1)
    raw = zeros(N_ν, ne)
    norm = zeros(N_ν, ne)
    check(ccall((:m1lya_edge_spectra, libm1lya), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Csize_t),
                ctx, raw, norm, length(raw)), ctx)
    return permutedims(raw), permutedims(norm)
end

"""Copies the device state back into the LyαArrays (SaveTimeStepHDF5, /Checkpoint)."""
function download_state!(Lyα)
    ctx = ctxof(Lyα)
    for (f, A) in ((F_J, Lyα.J_Lyα), (F_J_OLD, Lyα.J_Lyα_old), (F_H_R, Lyα.H_r_Lyα), (F_H_Z, Lyα.H_Z_Lyα),
                   (F_H_R_OLD, Lyα.H_r_Lyα_old), (F_H_Z_OLD, Lyα.H_Z_Lyα_old), (F_F_LYA_R, Lyα.f_Lyα_r),
                   (F_F_LYA_Z, Lyα.f_Lyα_Z), (F_DRHOU_R, Lyα.Δρu_Lyα_r), (F_DRHOU_Z, Lyα.Δρu_Lyα_Z))
        download!(ctx, f, A)
    end
    return Lyα
end

# --- src/UpdateLyaJandH.jl:107-124 ---
function AdvectLyα!(c̃Δt::Float64, Lyα, geometry, u_r_star, u_Z_star, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r,
                    ∂u_Z_∂Z, u_r_over_r, BC_r_left::Symbol=:Reflect, BC_r_right::Symbol=:Outflow,
                    BC_Z_bot::Symbol=:Reflect, BC_Z_top::Symbol=:Outflow)
    ctx = ctxof(Lyα)
    GC.@preserve u_r_star u_Z_star ∂u_r_∂r ∂u_r_∂Z ∂u_Z_∂r ∂u_Z_∂Z u_r_over_r check(
        ccall((:m1lya_advect_lya, libm1lya), Cint,
              (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
               Ptr{Float64}, Ptr{Float64}, Cint, Cint, Cint, Cint),
              ctx, c̃Δt, u_r_star, u_Z_star, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z, u_r_over_r,
              bc(BC_r_left), bc(BC_r_right), bc(BC_Z_bot), bc(BC_Z_top)), ctx)
    return nothing
end

# --- src/ConstantsAndFunctions.jl:483-500 (arrays are the LyαArrays members, already resident) ---
function LyαEmissionAbsorptionUpdate!(J_Lyα, J_Lyα_old, H_r_Lyα, H_Z_Lyα, H_r_old_Lyα, H_Z_old_Lyα,
                                      ∫dν_js_Lyα, ϕ_ν, Hvoigt, α_Lyα, αc_Lyα, α_F_Lyα, pd, Δρu_Lyα_r,
                                      Δρu_Lyα_Z, Δν, c̃::Float64, Δt::Float64)
    ctx = ctxof(only(k for (k, _) in CONTEXTS if k.J_Lyα === J_Lyα))
    check(ccall((:m1lya_emission_absorption_update, libm1lya), Cint, (Ptr{Cvoid}, Float64, Float64),
                ctx, c̃, Δt), ctx)
    return nothing
end

# --- src/UpdateLyaJandH.jl:231-244 ---
function FrequencyUpdateJLyα_AllCells!(c̃Δt::Float64, Lyα, T, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z,
                                       u_r_over_r, UpdateFreqFactors::Bool=true)
    ctx = ctxof(Lyα)
    GC.@preserve T ∂u_r_∂r ∂u_r_∂Z ∂u_Z_∂r ∂u_Z_∂Z u_r_over_r check(
        ccall((:m1lya_frequency_update_all_cells, libm1lya), Cint,
              (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
               Ptr{Float64}, Cint),
              ctx, c̃Δt, T, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z, u_r_over_r, UpdateFreqFactors), ctx)
    return nothing
end

# --- src/FreqFactors.jl:142-149 ---
function UpdateFreqFactors!(Lyα, T, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z, u_r_over_r)
    ctx = ctxof(Lyα)
    GC.@preserve T ∂u_r_∂r ∂u_r_∂Z ∂u_Z_∂r ∂u_Z_∂Z u_r_over_r check(
        ccall((:m1lya_update_freq_factors, libm1lya), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
              ctx, T, ∂u_r_∂r, ∂u_r_∂Z, ∂u_Z_∂r, ∂u_Z_∂Z, u_r_over_r), ctx)
    return nothing
end

# --- src/ConstantsAndFunctions.jl:588-599 (f_Lyα_r/Z are refreshed on the device) ---
function Update_f_Lyα_MF_L_Lyα!(f_Lyα_r, f_Lyα_Z, Δρu_Lyα_r, Δρu_Lyα_Z, ∫dν_js, ΔV, rCoM, ZCoM,
                                Δt::Float64, Z_center::Float64, BC_Z_bot::Symbol=:Reflect)
    ctx = ctxof(only(k for (k, _) in CONTEXTS if k.f_Lyα_r === f_Lyα_r))
    out = zeros(4)
    check(ccall((:m1lya_update_f_mf_l, libm1lya), Cint, (Ptr{Cvoid}, Float64, Float64, Cint, Ptr{Float64}),
                ctx, Δt, Z_center, bc(BC_Z_bot), out), ctx)
    return out[1], out[2], out[3], out[4]
end

# --- src/ConstantsAndFunctions.jl:691-704, the variant with gravity (f_grav_r/Z [N_r, N_Z]) ---
function Update_f_Lyα_MF_L_Lyα!(f_Lyα_r, f_Lyα_Z, Δρu_Lyα_r, Δρu_Lyα_Z, ∫dν_js, ΔV, rCoM, ZCoM,
                                f_grav_r::AbstractMatrix{Float64}, f_grav_Z::AbstractMatrix{Float64},
                                Δt::Float64, Z_center::Float64, BC_Z_bot::Symbol=:Reflect)
    ctx = ctxof(only(k for (k, _) in CONTEXTS if k.f_Lyα_r === f_Lyα_r))
    out = zeros(5)
    GC.@preserve f_grav_r f_grav_Z check(ccall((:m1lya_update_f_mf_l_gravity, libm1lya), Cint,
        (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Float64, Cint, Ptr{Float64}),
        ctx, Δt, f_grav_r, f_grav_Z, Z_center, bc(BC_Z_bot), out), ctx)
    return out[1], out[2], out[3], out[4], out[5]
end

# --- src/ConstantsAndFunctions.jl:969-976 ---
function Compute_f_esc_Lyα_Cylindrical(H_r_Lya, H_Z_Lya, A_r, A_Z, Δν, L_Lyα_allspace::Float64,
                                       BC_Z_bot::Symbol=:Reflect)
    ctx = ctxof(only(k for (k, _) in CONTEXTS if k.H_r_Lyα === H_r_Lya))
    f = Ref{Float64}(0.0)
    check(ccall((:m1lya_compute_f_esc, libm1lya), Cint, (Ptr{Cvoid}, Float64, Cint, Ref{Float64}),
                ctx, L_Lyα_allspace, bc(BC_Z_bot), f), ctx)
    return f[]
end

# --- src/ConstantsAndFunctions.jl:809-816 ---
function Reset_Δρu!(Δρu_r, Δρu_Z)
    ctx = ctxof(only(k for (k, _) in CONTEXTS if k.Δρu_Lyα_r === Δρu_r))
    check(ccall((:m1lya_reset_drhou, libm1lya), Cint, (Ptr{Cvoid},), ctx), ctx)
    return nothing
end

# --- src/ConstantsAndFunctions.jl:416-420 ---
function Update_e_Lyα!(e_Lyα, Δν, J)
    ctx = ctxof(only(k for (k, _) in CONTEXTS if k.J_Lyα === J))
    check(ccall((:m1lya_update_e_lya, libm1lya), Cint, (Ptr{Cvoid},), ctx), ctx)
    download!(ctx, F_E_LYA, e_Lyα)
    return nothing
end

# --- mixed-precision batch (m1lya_batch_*): many static clouds per launch ---

function bcheck(rc::Cint, b::Ptr{Cvoid})
    rc == 0 && return nothing
    buf = Vector{UInt8}(undef, 512)
    ccall((:m1lya_batch_last_error, libm1lya), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Csize_t), b, buf, length(buf))
    error("libm1lya batch status $rc: " * unsafe_string(pointer(buf)))
end

"""
    BatchM1Lya(clouds) -> Ptr{Cvoid}

`clouds` is a vector of named tuples `(Lyα, geometry, T, c̃, Δt, Z_center, BC_Z_bot)`,
one per static cloud as `SetupStaticCloudTest` builds it (src/StaticCloudTest.jl:92-315).
Returns the batch handle; step it with `run_batch!`.
"""
function BatchM1Lya(clouds)
    n = length(clouds)
    nn = Cint[size(c.Lyα.J_Lyα, 1) for c in clouds]
    nr = Cint[size(c.Lyα.J_Lyα, 2) for c in clouds]
    nz = Cint[size(c.Lyα.J_Lyα, 3) for c in clouds]
    h = Ref{Ptr{Cvoid}}(C_NULL)
    bcheck(ccall((:m1lya_batch_create, libm1lya), Cint, (Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ref{Ptr{Cvoid}}),
                 n, nn, nr, nz, h), C_NULL)
    b = h[]
    for (idx, c) in enumerate(clouds)
        i = Cint(idx - 1); Lyα = c.Lyα
        GC.@preserve Lyα c begin   # c.geometry, c.T and the Lyα arrays stay alive for the ccalls
            cg = cgeometry(c.geometry)
            bcheck(ccall((:m1lya_batch_set_geometry, libm1lya), Cint, (Ptr{Cvoid}, Cint, Ref{CGeometry}), b, i, cg), b)
            fg = cfreqgrid(Lyα)
            bcheck(ccall((:m1lya_batch_set_freq_grid, libm1lya), Cint, (Ptr{Cvoid}, Cint, Ref{CFreqGrid}), b, i, fg), b)
            J0 = max(maximum(Lyα.J_Lyα), 1e-300)
            bcheck(ccall((:m1lya_batch_set_scale, libm1lya), Cint, (Ptr{Cvoid}, Cint, Float64), b, i, J0), b)
            for (f, A) in ((F_J, Lyα.J_Lyα), (F_H_R, Lyα.H_r_Lyα), (F_H_Z, Lyα.H_Z_Lyα), (F_ALPHA_F, Lyα.α_F_Lyα),
                           (F_HVOIGT, Lyα.Hvoigt), (F_PHI_NU, Lyα.ϕ_ν), (F_ALPHA_LYA, Lyα.α_Lyα),
                           (F_ALPHA_C, Lyα.αc_Lyα), (F_PD, Lyα.pd), (F_JS, Lyα.∫dν_js_Lyα), (F_DNU_D, Lyα.ΔνD),
                           (F_T, c.T))
                bcheck(ccall((:m1lya_batch_upload, libm1lya), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Csize_t),
                             b, i, f, A, length(A)), b)
            end
        end
        bcheck(ccall((:m1lya_batch_set_step, libm1lya), Cint,
                     (Ptr{Cvoid}, Cint, Float64, Float64, Float64, Cint, Cint, Cint),
                     b, i, c.c̃, c.Δt, c.Z_center, bc(:Outflow), bc(c.BC_Z_bot), bc(:Outflow)), b)
    end
    return b
end

"""Runs `nsteps` steps of every cloud; returns the [8, nclouds, nsteps] step scalars
(M_F, M_F_radial_net, M_F_radial_out, L_Lyα, f_esc, kmin, kmax, 0)."""
function run_batch!(b::Ptr{Cvoid}, nsteps::Integer)
    n = ccall((:m1lya_batch_count, libm1lya), Cint, (Ptr{Cvoid},), b)
    hist = Array{Float64}(undef, 8, n, nsteps)
    bcheck(ccall((:m1lya_batch_run_steps, libm1lya), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), b, nsteps, hist), b)
    return hist
end

"""Steps every cloud until each one's ΔJ/J < tol, or max_steps; returns the [8, nclouds, steps]
step scalars (row 8 is ΔJ/J) and, per cloud, the step after which it was first below tol (0: never)."""
function run_batch_until_steady!(b::Ptr{Cvoid}, max_steps::Integer, tol::Float64)
    n = ccall((:m1lya_batch_count, libm1lya), Cint, (Ptr{Cvoid},), b)
    hist = Array{Float64}(undef, 8, n, max(max_steps, 1))
    steady = zeros(Cint, n)
    done = Ref{Cint}(0)
    bcheck(ccall((:m1lya_batch_run_until_steady, libm1lya), Cint,
                 (Ptr{Cvoid}, Cint, Float64, Ptr{Float64}, Ptr{Cint}, Ref{Cint}), b, max_steps, tol, hist, steady, done), b)
    return hist[:, :, 1:done[]], steady
end

# --- the M_F of each of n fused steps: the primal whose rrule (ext/M1LyaGPUChainRulesCoreExt.jl)
#     replays the steps through m1lya_adjoint_run ---
function run_MF(Lyα, nsteps::Integer, c̃::Float64, Δt::Float64, Z_center::Float64; BC_Z_bot::Symbol=:Reflect)
    ctx = ctxof(Lyα)
    hist = Array{Float64}(undef, 8, nsteps)
    check(ccall((:m1lya_run_steps, libm1lya), Cint, (Ptr{Cvoid}, Cint, Float64, Float64, Float64, Cint, Cint,
                Cint, Ptr{Float64}), ctx, nsteps, c̃, Δt, Z_center, bc(:Outflow), bc(BC_Z_bot), bc(:Outflow), hist), ctx)
    return hist[1, :]
end

# --- reverse-mode gradient of n fused steps (m1lya_adjoint_run) ---

"""
    adjoint_run!(Lyα, nsteps, c̃, Δt, Z_center; BC_Z_bot=:Reflect, mf_seed=nothing, spec_seed=nothing,
                 checkpoint_every=0)

Runs `nsteps` fused steps of the attached LyαArrays (as `m1lya_run_steps`) and returns
`(objective, history, grad_state, grad_cell, grad_bin)` for
Obj = Σ_s mf_seed[:, s] ⋅ (M_F, M_F_radial_net, M_F_radial_out)(s) + spec_seed ⋅ S(final), with
S the escaping spectrum (f_esc = Σ_k Δν_k S_k).  grad_state is [N_ν, N_r, N_Z, 3] (J, H_r, H_Z of
the initial state), grad_cell [N_r, N_Z, 5] (α_Lyα, α_c, ∫js, ΔνD, T-direct), grad_bin
[N_ν, N_r, N_Z, 3] (α_F, Hvoigt, ϕ_ν).  A ChainRules `rrule` around the step loop is this call.
"""
function adjoint_run!(Lyα, nsteps::Integer, c̃::Float64, Δt::Float64, Z_center::Float64;
                      BC_Z_bot::Symbol=:Reflect, mf_seed=nothing, spec_seed=nothing, checkpoint_every::Integer=0)
    ctx = ctxof(Lyα)
    N_ν, N_r, N_Z = size(Lyα.J_Lyα)
    mf = mf_seed === nothing ? C_NULL : convert(Matrix{Float64}, mf_seed)      # [3, nsteps]
    sp = spec_seed === nothing ? C_NULL : convert(Vector{Float64}, spec_seed)  # [N_ν]
    obj = Ref{Float64}(0.0)
    hist = Array{Float64}(undef, 8, nsteps)
    gs = Array{Float64}(undef, N_ν, N_r, N_Z, 3)
    gc = Array{Float64}(undef, N_r, N_Z, 5)
    gb = Array{Float64}(undef, N_ν, N_r, N_Z, 3)
    GC.@preserve mf sp check(ccall((:m1lya_adjoint_run, libm1lya), Cint,
        (Ptr{Cvoid}, Cint, Float64, Float64, Float64, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Cint,
         Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        ctx, nsteps, c̃, Δt, Z_center, bc(:Outflow), bc(BC_Z_bot), bc(:Outflow),
        mf === C_NULL ? Ptr{Float64}(C_NULL) : pointer(mf), sp === C_NULL ? Ptr{Float64}(C_NULL) : pointer(sp),
        checkpoint_every, obj, hist, gs, gc, gb), ctx)
    return obj[], hist, gs, gc, gb
end

"""
    set_adjoint_mode!(Lyα, mode)

0 (default): `adjoint_run!` differentiates each advection interface once in reverse mode;
1: dual-number passes per cell, the cross-check of mode 0 (same gradient to rounding).
"""
set_adjoint_mode!(Lyα, mode::Integer) =
    check(ccall((:m1lya_set_adjoint_mode, libm1lya), Cint, (Ptr{Cvoid}, Cint), ctxof(Lyα), mode), ctxof(Lyα))

# --- device-side coefficient refresh (m1lya_set_voigt_lut, m1lya_prepare_subcycling) ---

struct CLineConstants      # m1lya_line_constants
    nu_lya::Float64; b_prefact::Float64; nu_over_c::Float64; a_prefact::Float64; sigma0_prefact::Float64
end

"""Copies the Voigt tables of CreateVoigtLUTs (src/FastInterpolations.jl:26-80) to the device:
`luts` is [nx, 5] (erf, dawson, exp(-x²), x exp(-x²), (1-2x²) exp(-x²)) on [-xmax, xmax]."""
function set_voigt_lut!(Lyα, luts::Matrix{Float64}, xmax::Float64, lc::CLineConstants)
    ctx = ctxof(Lyα)
    GC.@preserve luts check(ccall((:m1lya_set_voigt_lut, libm1lya), Cint,
                                  (Ptr{Cvoid}, Ptr{Float64}, Cint, Float64, Ref{CLineConstants}),
                                  ctx, luts, size(luts, 1), xmax, lc), ctx)
    return nothing
end

# --- src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:2090-2155, for all cells at once ---
"""PrepareLyαSubCycling! on the device: Hvoigt, ϕ_ν (over all bins), ΔνD, α_Lyα and α_F from
T and n_HI; αc (absorption), αc_F (flux), js and pd come from the thermochemistry."""
function prepare_subcycling!(Lyα, T::Matrix{Float64}, nHI::Matrix{Float64}; αc=nothing, αc_F=nothing, js=nothing,
                             pd=nothing, all_bins::Bool=true)
    ctx = ctxof(Lyα)
    GC.@preserve T nHI αc αc_F js pd check(ccall((:m1lya_prepare_subcycling, libm1lya), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint),
        ctx, T, nHI, optr(αc), optr(αc_F), optr(js), optr(pd), Cint(all_bins)), ctx)
    return nothing
end

# --- PrepareLyαSubCycling! from the chemical state (PhotoThermoChemistry.jl:2086-2155): ∫dν js of
#     recombinations and collisions, p_d, the dust and gas continuum opacities, then the refresh above ---
"""`chem` is [N_r, N_Z, 8] = T, n_HI, n_H2, n_HII, n_HeI, n_e, n_OHx, n_CHx; `ρd` [3, N_r, N_Z] (or
nothing); `κ` = (κd_abs_Lyα, κd_F_Lyα) of silicate, graphite, PAH.  Needs `set_voigt_lut!`."""
function prepare_subcycling_chem!(Lyα, chem::Array{Float64,3}, ρd, κ::AbstractVector{Float64}; z::Float64=0.0,
                                  ots::Bool=true, destruction::Bool=true, all_bins::Bool=true)
    ctx = ctxof(Lyα)
    length(κ) == 6 || throw(ArgumentError("κ holds the 6 dust opacities"))
    κv = collect(κ)
    GC.@preserve chem ρd κv check(ccall((:m1lya_prepare_subcycling_chem, libm1lya), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Cint, Cint, Cint),
        ctx, chem, optr(ρd), κv, z, Cint(ots), Cint(destruction), Cint(all_bins)), ctx)
    return nothing
end

# --- steady-state option of the RT-only static cloud (the reference runs a fixed horizon,
#     src/StaticCloudTest.jl:317-322): fused steps until ΔJ/J < tol ---
"""Runs fused steps until ΔJ/J = Σ Δν ΔV |J_new - J_start| / Σ Δν ΔV max(J_new - Jeps, 0) < tol or
max_steps; returns the [8, steps] history (row 8 is ΔJ/J)."""
function run_until_steady!(Lyα, max_steps::Integer, tol::Float64, c̃::Float64, Δt::Float64, Z_center::Float64;
                           BC_r_right::Symbol=:Outflow, BC_Z_bot::Symbol=:Reflect, BC_Z_top::Symbol=:Outflow)
    ctx = ctxof(Lyα)
    hist = zeros(8, max(max_steps, 1))
    done = Ref{Cint}(0)
    check(ccall((:m1lya_run_until_steady, libm1lya), Cint,
                (Ptr{Cvoid}, Cint, Float64, Float64, Float64, Float64, Cint, Cint, Cint, Ptr{Float64}, Ref{Cint}),
                ctx, max_steps, tol, c̃, Δt, Z_center, bc(BC_r_right), bc(BC_Z_bot), bc(BC_Z_top), hist, done), ctx)
    return hist[:, 1:done[]]
end

# --- src/AdvectJandH.jl:29-68: the CFL step as a device reduction on the context of this grid ---
function GetRHDCFLTimeStep(r::AbstractVector{Float64}, Z::AbstractVector{Float64}, c̃::Float64,
                           u_r::AbstractArray{Float64,2}, u_Z::AbstractArray{Float64,2}, CFL::Float64)
    ctx = only(c for (k, c) in CONTEXTS if size(k.α_Lyα) == size(u_r))
    dt = Ref{Float64}(0.0)
    GC.@preserve u_r u_Z check(ccall((:m1lya_cfl_time_step, libm1lya), Cint,
        (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}, Float64, Ref{Float64}), ctx, c̃, u_r, u_Z, CFL, dt), ctx)
    return dt[]
end

struct CRslInputs          # m1lya_rsl_inputs
    u_r::Ptr{Float64}; u_z::Ptr{Float64}; gamma_ad::Ptr{Float64}; P::Ptr{Float64}; rho::Ptr{Float64}
    rho_d::Ptr{Float64}; xHII::Ptr{Float64}; xH2::Ptr{Float64}; T::Ptr{Float64}
    kappa_abs::NTuple{3,Float64}; lc::CLineConstants
end

# --- src/PhotoChemistryAndCooling/PhotoThermoChemistry.jl:616-722 ---
function FindRSL_LyαRT(c̃_Lyα::Float64, c̃_min::Float64, Lyα, Hydro, Chem, Dust, Geom; lc::CLineConstants)
    isnothing(Lyα) && return 0.0
    ctx = ctxof(Lyα)
    u_r, u_Z, P, ρ, ρd, γad, T = Hydro.u_X, Hydro.u_Y, Hydro.P, Hydro.ρ, Hydro.ρd, Hydro.γad, Hydro.T
    xHII, xH2 = Chem.xHII, Chem.xH2
    out = Ref{Float64}(0.0)
    GC.@preserve u_r u_Z P ρ ρd γad T xHII xH2 begin
        ri = CRslInputs(pointer(u_r), pointer(u_Z), pointer(γad), pointer(P), pointer(ρ), pointer(ρd),
                        pointer(xHII), pointer(xH2), pointer(T),
                        (Dust.κd_abs_Lyα_Sil, Dust.κd_abs_Lyα_C, Dust.κd_abs_Lyα_PAH), lc)
        check(ccall((:m1lya_find_rsl, libm1lya), Cint, (Ptr{Cvoid}, Float64, Float64, Ref{CRslInputs}, Ref{Float64}),
                    ctx, c̃_Lyα, c̃_min, ri, out), ctx)
    end
    return out[]
end

# --- the Lyα block of AbsorptionStep (PhotoThermoChemistry.jl:1577-1670) for every cell at once ---
"""Sub-step Δt_cells [N_r, N_Z] of every cell: updates J, H on the device from the *_old arrays,
rejects (and restores) cells whose e_Lyα changes by more than tol_Abs; returns (redo::Matrix{Cint},
out::Array{Float64,3}) with out[:, :, m] = Δe, e_Lyα_new, Δρu_r/Z_Lyα, then r/Z of gas, Sil, C, PAH."""
function absorption_substep!(Lyα, Δt_cells::Matrix{Float64}, c̃_Lyα::Float64, ρd, Dust)
    ctx = ctxof(Lyα)
    redo = zeros(Cint, size(Δt_cells))
    out = zeros(size(Δt_cells)..., 12)
    κF = [Dust.κd_F_Lyα_Sil, Dust.κd_F_Lyα_C, Dust.κd_F_Lyα_PAH]
    GC.@preserve Δt_cells ρd κF check(ccall((:m1lya_absorption_substep, libm1lya), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Float64, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Cint}, Ptr{Float64}),
        ctx, Δt_cells, 0.0, c̃_Lyα, ρd, κF, redo, out), ctx)
    return redo, out
end

end # module
