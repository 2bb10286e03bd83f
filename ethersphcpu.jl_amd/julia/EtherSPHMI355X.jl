#=
  EtherSPHMI355X.jl — Julia front end of libethersph_hip.so.

  Adds GPU methods of EtherSPHCPU.solve! selected by a trailing `MI355X()`
  argument; the particle types, kernels, models and time discretisation stay
  the reference's own.  Particle vectors are packed into contiguous
  Matrix{Float64}(dim, n) arrays (column-major == particle-major, the layout
  the C ABI expects), handed over once, stepped on the GPU and unpacked.

  Interfaces replaced (EtherSPHCPU.jl):
    solve!(fluid, wall, kernel, model, td, vtp_io; check)            Solve/FluidWallSolve.jl:85-140
    solve!(fluid, static_velocity_wall, ...)                          Solve/FluidVelocitySolve.jl:104-168
    solve!(fluid, thermostatic_wall, ..., ::ThermalWeakly...Model)    Solve/FluidThermostaticWallSolve.jl:127-173
    solve!(fluid, wall, thermostatic_wall, ..., ::ThermalWeakly...)   Solve/FluidCompulsiveThermostaticWallSolve.jl:154-233
    solve!(fluid, inlet, outlet, wall, ...; inlet/outlet closures)     Solve/FluidInletOutletWallSolve.jl:266-340
    kernelValue / kernelGradient                                      SPHKernel/SmoothKernel/*.jl
    CellListMap.neighborlist (as used by findNeighbours!)             Neighbour/Neighbour.jl:88
  `check` closures cannot run on the device; pass an AABBCheck instead.  The
  inlet/outlet closures are planes along an axis (InletOutletPlanes) and the
  examples' user_defined_function a named profile (BufferProfile).
=#
module EtherSPHMI355X

using EtherSPHCPU

export MI355X, AABBCheck, InletOutletPlanes, BufferProfile, GPUContext, gpu_neighborlist

const libesph = joinpath(@__DIR__, "..", "libethersph_hip.so")

struct MI355X
    device::Int32
end
MI355X() = MI355X(0)

"""The `check` predicate as data: keep iff lo <= x <= hi on every axis and
n . x <= d for every plane (n, d) (at most 8; a convex keep-region)."""
struct AABBCheck
    lo::Vector{Float64}
    hi::Vector{Float64}
    planes::Vector{Tuple{Vector{Float64},Float64}}
end
AABBCheck(lo, hi) = AABBCheck(lo, hi, Tuple{Vector{Float64},Float64}[])

"Arithmetic form: `:fp64` (the reference's) or `:fp32_bulk` (binary32 pair terms, binary64 integration)."
precision_code(p::Symbol) = p === :fp64 ? Int32(0) : p === :fp32_bulk ? Int32(1) :
    throw(ArgumentError("precision must be :fp64 or :fp32_bulk"))

"The inlet/outlet closures of FluidInletOutletWallSolve.jl:275-279 as planes along `axis` (1-based)."
struct InletOutletPlanes
    axis::Int
    inlet_enter::Float64     # inlet particle with x[axis] > inlet_enter joins the fluid...
    inlet_shift::Float64     # ...and moves back by inlet_shift
    outlet_enter::Float64    # fluid particle with x[axis] > outlet_enter joins the outlet
    outlet_exit::Float64     # outlet particle with x[axis] > outlet_exit is removed
end

"The examples' user_defined_function: 1 = 2D channel start-up series, 2 = 3D pipe profile."
struct BufferProfile
    kind::Int32
    f::Float64
    nu::Float64
    l::Float64
    center::NTuple{3,Float64}
    terms::Int32
    outlet_profile::Bool
end

# ---- C structures (include/ethersph.h) --------------------------------------
struct EsphSceneDesc
    dim::Int32
    kernel::Int32
    influence_radius::Float64
    solver::Int32
    precision::Int32
    rho_0::Float64
    c_0::Float64
    p_0::Float64
    gamma::Int32
    mu_0::Float64
    body_force::NTuple{3,Float64}
    t_0::Float64
    kappa_0::Float64
    cp::Float64
    beta::Float64
    dt::Float64
    density_reinitialized_interval::Int64
    check_interval::Int64
    check_enabled::Int32
    check_min::NTuple{3,Float64}
    check_max::NTuple{3,Float64}
    fluid_capacity::Int64
    domain_set::Int32
    domain_min::NTuple{3,Float64}
    domain_max::NTuple{3,Float64}
    io_axis::Int32
    inlet_enter::Float64
    inlet_shift::Float64
    outlet_enter::Float64
    outlet_exit::Float64
    profile::Int32
    profile_terms::Int32
    outlet_profile::Int32
    profile_f::Float64
    profile_nu::Float64
    profile_l::Float64
    profile_center::NTuple{3,Float64}
    verlet_skin::Float64
    check_nplanes::Int32
    check_planes::NTuple{32,Float64}   # 8 planes x (n_x, n_y, n_z, d)
end

struct EsphFluidArrays
    n::Int64
    x::Ptr{Float64}
    v::Ptr{Float64}
    rho::Ptr{Float64}
    p::Ptr{Float64}
    c::Ptr{Float64}
    mass::Ptr{Float64}
    gap::Ptr{Float64}
    t::Ptr{Float64}
    kappa::Ptr{Float64}
    cp::Ptr{Float64}
end

struct EsphWallArrays
    kind::Int32
    n::Int64
    x::Ptr{Float64}
    normal::Ptr{Float64}
    gap::Ptr{Float64}
    v::Ptr{Float64}
    t::Ptr{Float64}
    kappa::Ptr{Float64}
end

kernel_kind(::CubicSpline) = Int32(0)
kernel_kind(::Gaussian) = Int32(1)
kernel_kind(::WendlandC2) = Int32(2)
kernel_kind(::WendlandC4) = Int32(3)

wall_kind(::AbstractVector{<:CompulsiveWallParticle}) = Int32(0)
wall_kind(::AbstractVector{<:StaticVelocityWallParticle}) = Int32(1)
wall_kind(::AbstractVector{<:ThermostaticWallParticle}) = Int32(2)

# ---- context ---------------------------------------------------------------------
mutable struct GPUContext
    handle::Ptr{Cvoid}
    function GPUContext(device::Integer = 0, hbm_budget::Integer = 0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:esph_ctx_create, libesph), Cint, (Int32, UInt64, Ptr{Ptr{Cvoid}}), device, hbm_budget, h)
        rc == 0 || error("esph_ctx_create failed ($rc)")
        ctx = new(h[])
        finalizer(c -> ccall((:esph_ctx_destroy, libesph), Cint, (Ptr{Cvoid},), c.handle), ctx)
        return ctx
    end
end

function chk(ctx::GPUContext, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:esph_last_error, libesph), Cstring, (Ptr{Cvoid},), ctx.handle))
    error("ethersph error $rc: $msg")
end

const CONTEXTS = Dict{Int32,GPUContext}()
context(dev::MI355X) = get!(() -> GPUContext(dev.device), CONTEXTS, dev.device)

# ---- packing ------------------------------------------------------------------------
packvec(ps, f) = reduce(hcat, [getfield(p, f) for p in ps]; init = zeros(Float64, length(getfield(first(ps), f)), 0))
packsc(ps, f) = Float64[getfield(p, f) for p in ps]

function descriptor(dim, kernel, model, td, solver, check; planes = nothing, profile = nothing, capacity = 0,
                    precision::Symbol = :fp64)
    bf = ntuple(k -> k <= dim ? Float64(model.body_force_vec_[k]) : 0.0, 3)
    thermal = model isa ThermalWeaklyCompressibleLiquidModel
    lo = ntuple(k -> (check !== nothing && k <= dim) ? check.lo[k] : -Inf, 3)
    hi = ntuple(k -> (check !== nothing && k <= dim) ? check.hi[k] : Inf, 3)
    pl = planes === nothing ? InletOutletPlanes(1, 0.0, 0.0, 0.0, 0.0) : planes
    pr = profile === nothing ? BufferProfile(0, 0.0, 0.0, 0.0, (0.0, 0.0, 0.0), 0, false) : profile
    cpl = check === nothing ? Tuple{Vector{Float64},Float64}[] : check.planes
    length(cpl) <= 8 || throw(ArgumentError("at most 8 check planes"))
    flat = zeros(32)
    for (q, (nrm, d)) in enumerate(cpl)
        for k in 1:dim
            flat[4 * (q - 1) + k] = nrm[k]
        end
        flat[4 * (q - 1) + 4] = d
    end
    return EsphSceneDesc(dim, kernel_kind(kernel), kernel.influence_radius_, solver, precision_code(precision),
                         model.rho_0_, model.c_0_, model.p_0_, model.gamma_, model.mu_0_, bf,
                         thermal ? model.t_0_ : 0.0, thermal ? model.kappa_0_ : 0.0,
                         thermal ? model.cp_ : 0.0, thermal ? model.beta_ : 0.0,
                         td.dt_, td.density_reinitialized_interval_, td.check_interval_,
                         check === nothing ? 0 : 1, lo, hi, capacity, 0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0),
                         pl.axis - 1, pl.inlet_enter, pl.inlet_shift, pl.outlet_enter, pl.outlet_exit,
                         pr.kind, pr.terms, pr.outlet_profile ? 1 : 0, pr.f, pr.nu, pr.l, pr.center, 0.0,
                         length(cpl), ntuple(k -> flat[k], 32))
end

fluid_arrays(n, x, v, rho, p, c, m, gap) =
    EsphFluidArrays(n, pointer(x), pointer(v), pointer(rho), pointer(p), pointer(c), pointer(m), pointer(gap),
                    C_NULL, C_NULL, C_NULL)

pack_liquid(ps) = (packvec(ps, :x_vec_), packvec(ps, :v_vec_), packsc(ps, :rho_), packsc(ps, :p_), packsc(ps, :c_),
                   packsc(ps, :mass_), packsc(ps, :gap_))

# Reads the device state back into the caller's fluid vector in reference
# order (deleteat! of the particles the check removed, order kept).
function unpack_fluid!(ctx, scene, fluid, dim, thermal)
    n = Ref{Int64}(0)
    chk(ctx, ccall((:esph_scene_fluid_count, libesph), Cint, (Ptr{Cvoid}, Ptr{Int64}), scene, n))
    get(field, w) = begin
        out = Vector{Float64}(undef, max(n[] * w, 1))
        wr = Ref{Int64}(0)
        chk(ctx, ccall((:esph_scene_get_fluid, libesph), Cint,
                         (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64, Ptr{Int64}), scene, field, out, length(out), wr))
        out[1:wr[]]
    end
    ids = Int.(get(10, 1)) .+ 1          # upload indices of the live particles
    X = reshape(get(0, dim), dim, :); V = reshape(get(1, dim), dim, :)
    R = get(2, 1); P = get(3, 1); C = get(4, 1)
    T = thermal ? get(7, 1) : Float64[]
    return ids, X, V, R, P, C, T
end

# solve! of Solve/FluidWallSolve.jl:85-140 (and the other fixed-count
# solvers) on the device: upload once, writeVTP at step 0, then step to each
# output step (isOutputStep, FixedTimeIntervalDiscretization.jl:18-20) and
# write frame div(step, output_interval_) there with the reference's own
# writeVTP (:95, :136).  A failed chunk of steps (a non-finite position met on
# the device) takes the reference's catch path (:130-134): the error and the
# step are printed, the state is dumped with writeVTP(step, ...) and stepping
# continues.  The device reports such an error for a chunk, so the printed
# step is the last step of the chunk it was met in.
function run_gpu!(fluid, walls, kernel, model, td, solver, check, vtp_io, dev::MI355X; precision::Symbol = :fp64)
    ctx = context(dev)
    dim = length(first(fluid).x_vec_)
    desc = Ref(descriptor(dim, kernel, model, td, solver, check; precision = precision))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    chk(ctx, ccall((:esph_scene_create, libesph), Cint, (Ptr{Cvoid}, Ptr{EsphSceneDesc}, Ptr{Ptr{Cvoid}}),
                     ctx.handle, desc, h))
    scene = h[]
    thermal = first(fluid) isa ThermalLiquidParticle && model isa ThermalWeaklyCompressibleLiquidModel
    upload = copy(fluid)                  # the particle objects in upload order
    groups() = Any[fluid, walls...]
    function write_back!()
        ids, X, V, R, P, C, T = unpack_fluid!(ctx, scene, fluid, dim, thermal)
        keep = upload[ids]
        resize!(fluid, length(keep)); fluid .= keep
        for (i, q) in enumerate(fluid)
            q.x_vec_ .= X[:, i]; q.v_vec_ .= V[:, i]
            q.rho_ = R[i]; q.p_ = P[i]; q.c_ = C[i]
            thermal && (q.t_ = T[i])
        end
    end
    try
        # walls: all groups at once
        wpacked = [(wall_kind(w), length(w), packvec(w, :x_vec_), packvec(w, :normal_vec_), packsc(w, :gap_),
                    w isa AbstractVector{<:StaticVelocityWallParticle} ? packvec(w, :v_vec_) : zeros(dim, length(w)),
                    w isa AbstractVector{<:ThermostaticWallParticle} ? packsc(w, :t_) : zeros(length(w)),
                    w isa AbstractVector{<:ThermostaticWallParticle} ? packsc(w, :kappa_) : zeros(length(w)))
                   for w in walls]
        GC.@preserve wpacked begin
            wa = [EsphWallArrays(k, n, pointer(x), pointer(nn), pointer(g), pointer(v), pointer(t), pointer(kap))
                  for (k, n, x, nn, g, v, t, kap) in wpacked]
            chk(ctx, ccall((:esph_scene_set_walls, libesph), Cint, (Ptr{Cvoid}, Int32, Ptr{EsphWallArrays}),
                             scene, length(wa), wa))
        end
        x = packvec(fluid, :x_vec_); v = packvec(fluid, :v_vec_)
        rho = packsc(fluid, :rho_); p = packsc(fluid, :p_); c = packsc(fluid, :c_)
        m = packsc(fluid, :mass_); gap = packsc(fluid, :gap_)
        t = thermal ? packsc(fluid, :t_) : Float64[]
        kap = thermal ? packsc(fluid, :kappa_) : Float64[]
        cp = thermal ? packsc(fluid, :cp_) : Float64[]
        GC.@preserve x v rho p c m gap t kap cp begin
            fa = Ref(EsphFluidArrays(length(fluid), pointer(x), pointer(v), pointer(rho), pointer(p), pointer(c),
                                     pointer(m), pointer(gap), thermal ? pointer(t) : C_NULL,
                                     thermal ? pointer(kap) : C_NULL, thermal ? pointer(cp) : C_NULL))
            chk(ctx, ccall((:esph_scene_set_fluid, libesph), Cint, (Ptr{Cvoid}, Ptr{EsphFluidArrays}), scene, fa))
        end
        assureDirPathExist(vtp_io)
        writeVTP(0, 0.0, vtp_io, groups())
        total = td.total_step_
        every = td.output_interval_ > 0 ? td.output_interval_ : total
        step = 0
        while step < total
            next = min(total, (div(step, every) + 1) * every)
            rc = ccall((:esph_scene_step, libesph), Cint, (Ptr{Cvoid}, Int64), scene, next - step)
            rc == 0 && (rc = ccall((:esph_scene_sync, libesph), Cint, (Ptr{Cvoid},), scene))
            step = next
            if rc != 0
                println(ErrorException("ethersph error $rc: " *
                                       unsafe_string(ccall((:esph_last_error, libesph), Cstring, (Ptr{Cvoid},),
                                                           ctx.handle))))
                println("Error at step: ", step)
                write_back!()
                writeVTP(step, step * td.dt_, vtp_io, groups())
            end
            if isOutputStep(step, td)
                write_back!()
                writeVTP(div(step, td.output_interval_), step * td.dt_, vtp_io, groups())
            end
        end
        write_back!()
    finally
        ccall((:esph_scene_destroy, libesph), Cint, (Ptr{Cvoid},), scene)
    end
    return nothing
end

# ---- solve! methods (dispatch mirrors Solve/*.jl) ------------------------------------
function EtherSPHCPU.solve!(fluid::AbstractVector{<:FluidParticle}, wall::AbstractVector{<:WallParticle},
                            kernel::SmoothKernel, model::WeaklyCompressibleLiquidModel,
                            td::DensityReinitializedFixedTimeIntervalForwardEuler, vtp_io, dev::MI355X;
                            check::Union{Nothing,AABBCheck} = nothing, precision::Symbol = :fp64)
    run_gpu!(fluid, [wall], kernel, model, td, Int32(0), check, vtp_io, dev; precision = precision)
end

function EtherSPHCPU.solve!(fluid::AbstractVector{<:FluidParticle}, wall::AbstractVector{<:StaticVelocityWallParticle},
                            kernel::SmoothKernel, model::WeaklyCompressibleLiquidModel,
                            td::DensityReinitializedFixedTimeIntervalForwardEuler, vtp_io, dev::MI355X;
                            check::Union{Nothing,AABBCheck} = nothing, precision::Symbol = :fp64)
    run_gpu!(fluid, [wall], kernel, model, td, Int32(1), check, vtp_io, dev; precision = precision)
end

function EtherSPHCPU.solve!(fluid::AbstractVector{<:FluidParticle}, wall::AbstractVector{<:ThermostaticWallParticle},
                            kernel::SmoothKernel, model::ThermalWeaklyCompressibleLiquidModel,
                            td::DensityReinitializedFixedTimeIntervalForwardEuler, vtp_io, dev::MI355X;
                            check::Union{Nothing,AABBCheck} = nothing, precision::Symbol = :fp64)
    run_gpu!(fluid, [wall], kernel, model, td, Int32(2), check, vtp_io, dev; precision = precision)
end

function EtherSPHCPU.solve!(fluid::AbstractVector{<:FluidParticle}, wall::AbstractVector{<:WallParticle},
                            thermo::AbstractVector{<:ThermostaticWallParticle}, kernel::SmoothKernel,
                            model::ThermalWeaklyCompressibleLiquidModel,
                            td::DensityReinitializedFixedTimeIntervalForwardEuler, vtp_io, dev::MI355X;
                            check::Union{Nothing,AABBCheck} = nothing, precision::Symbol = :fp64)
    run_gpu!(fluid, [wall, thermo], kernel, model, td, Int32(3), check, vtp_io, dev; precision = precision)
end

# FluidInletOutletWallSolve.jl:266-340: the groups change size (inlet copies join the
# fluid, leavers join the outlet, expired outlet particles are removed); the
# caller's vectors are refilled in reference order.
function EtherSPHCPU.solve!(fluid::AbstractVector{<:FluidParticle}, inlet::AbstractVector{<:FluidParticle},
                            outlet::AbstractVector{<:FluidParticle}, wall::AbstractVector{<:WallParticle},
                            kernel::SmoothKernel, model::WeaklyCompressibleLiquidModel,
                            td::DensityReinitializedFixedTimeIntervalForwardEuler, vtp_io, dev::MI355X;
                            inlet_outlet::InletOutletPlanes, user_defined_function::Union{Nothing,BufferProfile} = nothing,
                            precision::Symbol = :fp64)
    ctx = context(dev)
    dim = length(first(fluid).x_vec_)
    desc = Ref(descriptor(dim, kernel, model, td, Int32(4), nothing; planes = inlet_outlet,
                          profile = user_defined_function, precision = precision))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    chk(ctx, ccall((:esph_scene_create, libesph), Cint, (Ptr{Cvoid}, Ptr{EsphSceneDesc}, Ptr{Ptr{Cvoid}}),
                     ctx.handle, desc, h))
    scene = h[]
    protos = (deepcopy(first(fluid)), deepcopy(isempty(inlet) ? first(fluid) : first(inlet)),
              deepcopy(isempty(outlet) ? first(fluid) : first(outlet)))
    try
        wx = packvec(wall, :x_vec_); wn = packvec(wall, :normal_vec_); wg = packsc(wall, :gap_)
        wz = zeros(dim, length(wall)); wz1 = zeros(length(wall))
        GC.@preserve wx wn wg wz wz1 begin
            wa = [EsphWallArrays(wall_kind(wall), length(wall), pointer(wx), pointer(wn), pointer(wg), pointer(wz),
                                 pointer(wz1), pointer(wz1))]
            chk(ctx, ccall((:esph_scene_set_walls, libesph), Cint, (Ptr{Cvoid}, Int32, Ptr{EsphWallArrays}),
                             scene, 1, wa))
        end
        bi = pack_liquid(inlet); bo = pack_liquid(outlet); fl = pack_liquid(fluid)
        GC.@preserve bi bo fl begin
            ia = Ref(fluid_arrays(length(inlet), bi...)); oa = Ref(fluid_arrays(length(outlet), bo...))
            chk(ctx, ccall((:esph_scene_set_buffers, libesph), Cint,
                             (Ptr{Cvoid}, Ptr{EsphFluidArrays}, Ptr{EsphFluidArrays}), scene, ia, oa))
            fa = Ref(fluid_arrays(length(fluid), fl...))
            chk(ctx, ccall((:esph_scene_set_fluid, libesph), Cint, (Ptr{Cvoid}, Ptr{EsphFluidArrays}), scene, fa))
        end
        function write_back!()
            for (group, ps) in ((0, fluid), (1, inlet), (2, outlet))
                n = Ref{Int64}(0)
                chk(ctx, ccall((:esph_scene_group_count, libesph), Cint, (Ptr{Cvoid}, Int32, Ptr{Int64}),
                                 scene, group, n))
                get(field, w) = begin
                    out = Vector{Float64}(undef, max(n[] * w, 1))
                    wr = Ref{Int64}(0)
                    chk(ctx, ccall((:esph_scene_get_group, libesph), Cint,
                                     (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Int64, Ptr{Int64}),
                                     scene, group, field, out, length(out), wr))
                    out[1:wr[]]
                end
                proto = protos[group + 1]
                while length(ps) < n[]
                    push!(ps, deepcopy(proto))
                end
                resize!(ps, n[])
                X = reshape(get(0, dim), dim, :); V = reshape(get(1, dim), dim, :)
                R = get(2, 1); P = get(3, 1); C = get(4, 1); M = get(5, 1); G = get(6, 1)
                for (i, q) in enumerate(ps)
                    q.x_vec_ .= X[:, i]; q.v_vec_ .= V[:, i]
                    q.rho_ = R[i]; q.p_ = P[i]; q.c_ = C[i]; q.mass_ = M[i]; q.gap_ = G[i]
                end
            end
        end
        groups() = Any[fluid, inlet, outlet, wall]
        # as the fixed-count solvers: frame 0, then every output step
        # (FluidInletOutletWallSolve.jl:282,331)
        assureDirPathExist(vtp_io)
        writeVTP(0, 0.0, vtp_io, groups())
        total = td.total_step_
        every = td.output_interval_ > 0 ? td.output_interval_ : total
        step = 0
        while step < total
            next = min(total, (div(step, every) + 1) * every)
            chk(ctx, ccall((:esph_scene_step, libesph), Cint, (Ptr{Cvoid}, Int64), scene, next - step))
            chk(ctx, ccall((:esph_scene_sync, libesph), Cint, (Ptr{Cvoid},), scene))
            step = next
            if isOutputStep(step, td)
                write_back!()
                writeVTP(div(step, td.output_interval_), step * td.dt_, vtp_io, groups())
            end
        end
        write_back!()
    finally
        ccall((:esph_scene_destroy, libesph), Cint, (Ptr{Cvoid},), scene)
    end
    return nothing
end

# ---- kernels and neighbour lists on the device -----------------------------------------
function device_kernel(r::AbstractVector{Float64}, k::SmoothKernel; dev::MI355X = MI355X())
    ctx = context(dev)
    w = similar(r); dw = similar(r)
    chk(ctx, ccall((:esph_kernel_eval, libesph), Cint,
                     (Ptr{Cvoid}, Int32, Int32, Float64, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                     ctx.handle, kernel_kind(k), k.dim_, k.influence_radius_, length(r), r, w, dw))
    return w, dw
end

"Pairs (i, j) (1-based) within `cutoff`: within x (i < j), or between x and y."
function gpu_neighborlist(x::Matrix{Float64}, cutoff::Float64, y::Union{Nothing,Matrix{Float64}} = nothing;
                          dev::MI355X = MI355X())
    ctx = context(dev)
    dim = size(x, 1)
    ny = y === nothing ? 0 : size(y, 2)
    yp = y === nothing ? Ptr{Float64}(C_NULL) : pointer(y)
    m = Ref{Int64}(0)
    GC.@preserve x y begin
        chk(ctx, ccall((:esph_neighbor_list, libesph), Cint,
                         (Ptr{Cvoid}, Int32, Float64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Int32}, Int64, Ptr{Int64}),
                         ctx.handle, dim, cutoff, size(x, 2), x, ny, yp, C_NULL, 0, m))
        out = Matrix{Int32}(undef, 2, m[])
        chk(ctx, ccall((:esph_neighbor_list, libesph), Cint,
                         (Ptr{Cvoid}, Int32, Float64, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Int32}, Int64, Ptr{Int64}),
                         ctx.handle, dim, cutoff, size(x, 2), x, ny, yp, out, m[], m))
    end
    return [(Int(out[1, q]) + 1, Int(out[2, q]) + 1) for q in 1:m[]]
end

end # module
