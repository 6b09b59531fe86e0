# collop_hip.jl -- drop-in replacement of collision_gpu.jl + projection_gpu.jl
# for junyixu/CollisionOperators.jl: the same GPU hook functions, served by the
# gfx950 library libcollop.so through ccall (C-ABI: include/collop.h).
#
# Use: in main.jl's enable_gpu! (main.jl:390-423) include this file instead of
# collision_gpu.jl / projection_gpu.jl; the hook assignments there stay as they
# are, because the names below are the ones it looks up with getglobal.
# No CUDA.jl is loaded.  Set ENV["COLLOP_LIB"] to the library path if it is
# not at the default location next to this file.
#
# (Julia is not part of the build image this library was developed in, so this
# file is exercised only through the equivalent Python ctypes binding,
# collisionoperators.jl_amd/collop/hooks.py, which calls the same entry points
# with the same arguments.)

const _COLLOP_LIB = get(ENV, "COLLOP_LIB", joinpath(@__DIR__, "..", "collop", "libcollop.so"))
const _COLLOP_CTX = Ref{Ptr{Cvoid}}(C_NULL)
const _COLLOP_MESH = Ref{UInt}(0)          # objectid of the Workspace whose tables are loaded

function _collop_check(st::Cint)
    st == 0 && return nothing
    msg = unsafe_string(ccall((:collop_last_error, _COLLOP_LIB), Cstring, (Ptr{Cvoid},), _COLLOP_CTX[]))
    error("libcollop status $st: $msg")
end

function _collop_ctx()
    if _COLLOP_CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:collop_create, _COLLOP_LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h)
        st == 0 || error("collop_create failed (status $st): no usable HIP device")
        _COLLOP_CTX[] = h[]
        atexit(() -> ccall((:collop_destroy, _COLLOP_LIB), Cint, (Ptr{Cvoid},), _COLLOP_CTX[]))
    end
    return _COLLOP_CTX[]
end

# Hand the Workspace tables to the library once per Workspace (the reference
# caches the same tables on the device, projection_gpu.jl:186-211).
function _collop_mesh!(ws)
    ctx = _collop_ctx()
    id = objectid(ws)
    _COLLOP_MESH[] == id && return ctx
    ws.p.P_DEG == 2 || error("libcollop is specialized to P_DEG=2")
    flat(exts) = begin
        A = Array{Float64, 3}(undef, 3, 3, length(exts))
        for (k, C) in enumerate(exts)
            A[:, :, k] .= Matrix(C)
        end
        A
    end
    C1 = flat(ws.ext1d_1)
    C2 = flat(ws.ext1d_2)
    bs1 = Int32.(ws.basis_start_1d_1)
    bs2 = Int32.(ws.basis_start_1d_2)
    _collop_check(ccall((:collop_set_mesh, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64},
         Ptr{Int32}, Ptr{Int32}, Cint, Cint, Cint, Cint),
        ctx, ws.bp1, length(ws.bp1), ws.bp2, length(ws.bp2), C1, C2, bs1, bs2,
        ws.n_dofs_1d_1, ws.n_dofs_1d_2, ws.p.P_DEG, ws.p.N_QUAD))
    _COLLOP_MESH[] = id
    return ctx
end

_colmajor(A) = A isa Matrix{Float64} ? A : Matrix{Float64}(A)

# COLLISION_FN (solver.jl:30); replaces collision_gpu.jl:258-280
function compute_collision_gpu!(ws, dot_v, v_parts, w_parts, G)
    ctx = _collop_mesh!(ws)
    N = size(v_parts, 1)
    out = dot_v isa Matrix{Float64} ? dot_v : zeros(N, 2)
    _collop_check(ccall((:collop_collision_f64, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64),
        ctx, _colmajor(v_parts), Vector{Float64}(w_parts), _colmajor(G), out, N))
    out === dot_v || (dot_v .= out)
    return nothing
end

# COLLISION_FN with gpu_fp32 = true; replaces collision_gpu.jl:233-256
function compute_collision_gpu32!(ws, dot_v, v_parts, w_parts, G)
    ctx = _collop_mesh!(ws)
    N = size(v_parts, 1)
    out = dot_v isa Matrix{Float64} ? dot_v : zeros(N, 2)
    _collop_check(ccall((:collop_collision_f32c, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64),
        ctx, _colmajor(v_parts), Vector{Float64}(w_parts), _colmajor(G), out, N))
    out === dot_v || (dot_v .= out)
    return nothing
end

# COLLMETRIC_FN (solver.jl:34); replaces collision_gpu.jl:377-402
function compute_collision_metric_gpu!(ws, dot_v, A, v_parts, w_parts, G)
    ctx = _collop_mesh!(ws)
    N = size(v_parts, 1)
    F = zeros(N, 2)
    M = zeros(N, 3)
    _collop_check(ccall((:collop_collision_metric_f64, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64),
        ctx, _colmajor(v_parts), Vector{Float64}(w_parts), _colmajor(G), F, M, N))
    dot_v .= F
    A .= M
    return nothing
end

# L2PROJ_FN (solver.jl:31); replaces projection_gpu.jl:213-226.  The mass solve
# runs on the device; for bitwise host-solve parity use collop_l2_rhs and
# `ws.M_lu \ rhs` instead.
function l2_project_gpu!(ws, f_coeffs, v_parts, w_parts)
    ctx = _collop_mesh!(ws)
    out = Vector{Float64}(undef, ws.n_dofs)
    _collop_check(ccall((:collop_l2_project, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}),
        ctx, _colmajor(v_parts), Vector{Float64}(w_parts), size(v_parts, 1), out))
    f_coeffs .= out
    return nothing
end

# COMPG_FN (solver.jl:32); replaces projection_gpu.jl:228-246
function compute_G_gpu!(ws, G, v_parts, L_vec)
    ctx = _collop_mesh!(ws)
    N = size(v_parts, 1)
    out = zeros(N, 2)
    _collop_check(ccall((:collop_grad_gather, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}),
        ctx, _colmajor(v_parts), N, Vector{Float64}(L_vec), out))
    G .= out
    return nothing
end

# LOGGRAD_FN (solver.jl:33); replaces projection_gpu.jl:251-269 (FS_FLOOR, G_MAX
# of functions.jl:23,31)
function eval_loggrad_gpu!(ws, g, v_parts, f_coeffs)
    ctx = _collop_mesh!(ws)
    N = size(v_parts, 1)
    out = zeros(N, 2)
    _collop_check(ccall((:collop_loggrad_gather, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Float64, Float64, Ptr{Float64}),
        ctx, _colmajor(v_parts), N, Vector{Float64}(f_coeffs), 1e-30, 100.0, out))
    g .= out
    return nothing
end

# Run-record diagnostics (main.jl:306-307): device versions of
# compute_negative_part_l1 (functions.jl:300-313) and compute_fs_minus_fp_l2
# (functions.jl:320-343), taking the coefficient vector the field is built from.
function compute_negative_part_l1_gpu(ws, f_coeffs)
    ctx = _collop_mesh!(ws)
    out = Ref{Float64}(0.0)
    _collop_check(ccall((:collop_negative_part_l1, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ref{Float64}), ctx, Vector{Float64}(f_coeffs), out))
    return out[]
end

function compute_fs_minus_fp_l2_gpu(ws, f_coeffs, v_parts, w_parts)
    ctx = _collop_mesh!(ws)
    out = Ref{Float64}(0.0)
    _collop_check(ccall((:collop_fs_minus_fp_l2, _COLLOP_LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Ref{Float64}),
        ctx, _colmajor(v_parts), Vector{Float64}(w_parts), size(v_parts, 1), Vector{Float64}(f_coeffs), out))
    return out[]
end
