# RBFFDAmd.jl -- Julia binding of librbffd (include/rbffd.h) with the API of
# RadialBasisFiniteDifferences.jl's operator path:
#
#   generate_operator(X, Y, p, n, polydeg)                       src/generate_operator.jl:29
#   generate_operator(X, Y, p, n, polydeg, X_idx_in, X_idx_bc, X_idx_bc_g,
#                     Y_idx_in, Y_idx_bc, Y_idx_bc_g)            src/generate_operator.jl:192-193
#   generate_operator(X, p, n, polydeg)                          src/generate_operator.jl:354
#   hyperviscosity_operator(k, X, Y, p, n, polydeg[, six index sets])
#                                                                src/hyperviscosity_operator.jl:26,177-179
#   hyperviscosity_operator(k, X, p, n, polydeg)                 src/hyperviscosity_operator.jl:314
#
#   solve(D, f)  (= D \ f of test/poisson_test.jl:121)           least squares on the GPU
#   DeviceOperator(A); mul!(y, op, x); op * x                    A * x of test/poisson_test.jl:121,
#                                                                examples/adv_diff_test.jl:151-152
#
# Each generator returns SparseMatrixCSC{Float64,Int64} in the reference's layout:
# 1-based, rows sorted inside each column, explicit zeros kept, dims
# (number of Y nodes, largest referenced column).  Errors map to the
# exceptions the reference raises (SingularException from inv, ArgumentError,
# UndefRefError for a node in no index set).
#
# Set ENV["RBFFD_LIB"] to the path of librbffd.so.  One context per Julia
# thread; calls are synchronous.
module RBFFDAmd

using LinearAlgebra
using SparseArrays
using StaticArrays

export generate_operator, hyperviscosity_operator, solve, DeviceOperator

const LIB = get(ENV, "RBFFD_LIB", joinpath(@__DIR__, "..", "rbffd_amd", "librbffd.so"))

# operator kinds and precisions (include/rbffd.h)
const OP_E, OP_DX, OP_DY, OP_DXX, OP_DYY, OP_DXY, OP_XK, OP_YK = Int32.(0:7)
const PREC_FP64 = Int32(0)

# struct rbffd_boundary (same field order and C alignment)
struct Boundary
    index_base::Int32
    n_in::Int64
    in_idx::Ptr{Int64}
    n_markers::Int32
    bc_ptr::Ptr{Int64}
    bc_idx::Ptr{Int64}
    ghost_ptr::Ptr{Int64}
    ghost_idx::Ptr{Int64}
end

# struct rbffd_stats
struct Stats
    rows::Int64
    singular::Int64
    nonfinite::Int64
    refined_fallback::Int64
    ms_knn::Float64
    ms_solve::Float64
    ms_fallback::Float64
    lanes_per_stencil::Int64
    factorizations::Int64
end

const CONTEXTS = Dict{Int,Ptr{Cvoid}}()
const CONTEXT_LOCK = ReentrantLock()

function context()
    tid = Threads.threadid()
    lock(CONTEXT_LOCK) do
        get!(CONTEXTS, tid) do
            h = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:rbffd_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h))
            h[]
        end
    end
end

function check(code::Integer)
    code == 0 && return nothing
    msg = unsafe_string(ccall((:rbffd_last_error, LIB), Cstring, ()))
    code == 4 && throw(LinearAlgebra.SingularException(0))   # inv, src/interpolationmatrix.jl:8
    code == 6 && throw(UndefRefError())                      # src/calculateneighbors.jl:8-41
    (code == 1 || code == 2) && throw(ArgumentError(msg))
    error("librbffd: " * msg)
end

# index sets -> concatenated Int64 arrays with 0-based offsets
function _flatten(sets)
    ptr = Int64[0]
    idx = Int64[]
    for s in sets
        append!(idx, Int64.(collect(s)))
        push!(ptr, length(idx))
    end
    return ptr, idx
end

function _operators(X::AbstractVector{<:SVector{2,Float64}}, Y, p, n, polydeg,
                    ops::Vector{Int32}, k::Integer, sets)
    N = length(X)
    M = Y === nothing ? N : length(Y)
    nnz = M * n
    colptr = Vector{Int64}(undef, N + 1)
    rowval = Vector{Int64}(undef, nnz)
    nz = [Vector{Float64}(undef, nnz) for _ in ops]
    opk = fill(Int32(k), length(ops))
    ncols = Ref{Int64}(0)
    stats = Ref{Stats}()
    Xc = Vector{SVector{2,Float64}}(X)
    Yc = Y === nothing ? nothing : Vector{SVector{2,Float64}}(Y)
    in_idx = Int64[]; bc_ptr = Int64[]; bc_idx = Int64[]; g_ptr = Int64[]; g_idx = Int64[]
    if sets !== nothing
        X_idx_in, X_idx_bc, X_idx_bc_g = sets
        in_idx = Int64.(collect(X_idx_in))
        bc_ptr, bc_idx = _flatten(X_idx_bc)
        g_ptr, g_idx = _flatten(X_idx_bc_g)
    end
    code = GC.@preserve Xc Yc colptr rowval nz opk ops in_idx bc_ptr bc_idx g_ptr g_idx begin
        xp = Ptr{Float64}(pointer(Xc))
        yp = Yc === nothing ? Ptr{Float64}(C_NULL) : Ptr{Float64}(pointer(Yc))
        bnd = sets === nothing ? nothing :
              Ref(Boundary(Int32(1), length(in_idx), pointer(in_idx), Int32(length(bc_ptr) - 1),
                           pointer(bc_ptr), pointer(bc_idx), pointer(g_ptr), pointer(g_idx)))
        nzp = [pointer(v) for v in nz]
        GC.@preserve bnd nzp ccall((:rbffd_generate_operator, LIB), Cint,
              (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int32, Int32, Int32,
               Ptr{Boundary}, Int32, Ptr{Int32}, Ptr{Int32}, Int32, Int32,
               Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Float64}}, Ref{Int64}, Ref{Stats}),
              context(), xp, N, yp, M, p, n, polydeg,
              bnd === nothing ? Ptr{Boundary}(C_NULL) : Base.unsafe_convert(Ptr{Boundary}, bnd),
              length(ops), ops, opk, PREC_FP64, Int32(1),
              colptr, rowval, nzp, ncols, stats)
    end
    check(code)
    nc = ncols[]
    cp = colptr[1:nc+1]
    # every matrix owns its colptr and rowval, as the reference's separate
    # sparse() calls return independent matrices (src/generate_operator.jl:176-181)
    return Tuple(SparseMatrixCSC(M, nc, i == 1 ? cp : copy(cp), i == 1 ? rowval : copy(rowval), v)
                 for (i, v) in enumerate(nz))
end

const SIX = Int32[OP_E, OP_DX, OP_DY, OP_DXX, OP_DYY, OP_DXY]

generate_operator(X, Y, p, n, polydeg) = _operators(X, Y, p, n, polydeg, SIX, 0, nothing)

generate_operator(X, Y, p, n, polydeg, X_idx_in, X_idx_bc, X_idx_bc_g, Y_idx_in, Y_idx_bc, Y_idx_bc_g) =
    _operators(X, Y, p, n, polydeg, SIX, 0, (X_idx_in, X_idx_bc, X_idx_bc_g))

generate_operator(X, p, n, polydeg) = _operators(X, nothing, p, n, polydeg, SIX, 0, nothing)

hyperviscosity_operator(k_deriv, X, Y, p, n, polydeg) =
    _operators(X, Y, p, n, polydeg, Int32[OP_XK, OP_YK], k_deriv, nothing)

hyperviscosity_operator(k_deriv, X, Y, p, n, polydeg, X_idx_in, X_idx_bc, X_idx_bc_g,
                        Y_idx_in, Y_idx_bc, Y_idx_bc_g) =
    _operators(X, Y, p, n, polydeg, Int32[OP_XK, OP_YK], k_deriv, (X_idx_in, X_idx_bc, X_idx_bc_g))

hyperviscosity_operator(k_deriv, X, p, n, polydeg) =
    _operators(X, nothing, p, n, polydeg, Int32[OP_XK, OP_YK], k_deriv, nothing)

# ---- D \ f of the PDE harness (test/poisson_test.jl:121), rbffd_lsq_solve_device
const SOLVE_METHODS = Dict(:auto => Int32(0), :cgls => Int32(1), :qr => Int32(2))

function _dalloc(c, bytes)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:rbffd_device_malloc, LIB), Cint, (Ptr{Cvoid}, Int64, Ref{Ptr{Cvoid}}), c, max(bytes, 16), r))
    return r[]
end

_dfree(c, p) = ccall((:rbffd_device_free, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), c, p)

function _upload(c, a::Vector)
    p = _dalloc(c, sizeof(a))
    GC.@preserve a check(ccall((:rbffd_memcpy_h2d, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                               c, p, pointer(a), sizeof(a)))
    return p
end

# ---- A * x of the harnesses (test/poisson_test.jl:121, examples/adv_diff_test.jl:151-152)
# on the device: the CSR of A (the CSC of transpose(A)) is uploaded once and
# applied by rbffd_spmv_device; x and y cross PCIe per product.  Fields: a
# vector, or an N x F matrix with F in 1, 2, 4, 8 (node-major on the device).
mutable struct DeviceOperator
    m::Int
    n::Int
    ctx::Ptr{Cvoid}
    rowptr::Ptr{Cvoid}   # Int64, 0-based
    col::Ptr{Cvoid}      # Int32, 0-based
    vals::Ptr{Cvoid}     # Float64
    function DeviceOperator(A::SparseMatrixCSC{Float64})
        M, N = size(A)
        At = sparse(transpose(A))              # its colptr/rowval are the CSR of A
        c = context()
        rp = _upload(c, Int64.(At.colptr .- 1))
        cl = _upload(c, Int32.(At.rowval .- 1))
        vl = _upload(c, Vector{Float64}(At.nzval))
        op = new(M, N, c, rp, cl, vl)
        finalizer(o -> foreach(p -> _dfree(o.ctx, p), (o.rowptr, o.col, o.vals)), op)
        return op
    end
end

Base.size(op::DeviceOperator) = (op.m, op.n)

function LinearAlgebra.mul!(y::AbstractVecOrMat{Float64}, op::DeviceOperator, x::AbstractVecOrMat{Float64})
    F = size(x, 2)
    size(x, 1) == op.n && size(y, 1) == op.m && size(y, 2) == F || throw(DimensionMismatch("mul!"))
    xr = Matrix{Float64}(transpose(reshape(x, op.n, F)))  # node-major: x[f, node]
    yr = Matrix{Float64}(undef, F, op.m)
    c = op.ctx
    xd = _upload(c, vec(xr))
    yd = _dalloc(c, sizeof(yr))
    try
        check(ccall((:rbffd_spmv_device, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Ptr{Cvoid}, Ptr{Cvoid},
                     Float64, Float64, Int32, Ptr{Cvoid}),
                    c, op.m, op.rowptr, op.col, op.vals, Int32(0), Int32(F), xd, yd, 1.0, 0.0, Int32(0), C_NULL))
        GC.@preserve yr check(ccall((:rbffd_memcpy_d2h, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                                    c, pointer(yr), yd, sizeof(yr)))
    finally
        _dfree(c, xd)
        _dfree(c, yd)
    end
    y .= reshape(transpose(yr), size(y))
    return y
end

Base.:*(op::DeviceOperator, x::AbstractVector{Float64}) = mul!(Vector{Float64}(undef, op.m), op, x)
Base.:*(op::DeviceOperator, x::AbstractMatrix{Float64}) = mul!(Matrix{Float64}(undef, op.m, size(x, 2)), op, x)

function solve(D::SparseMatrixCSC{Float64}, f::AbstractVector{Float64}; method::Symbol = :auto,
               tol::Float64 = 1e-12, max_iters::Integer = 100_000)
    M, N = size(D)
    Dt = sparse(transpose(D))                  # the CSC of D' is the CSR of D
    c = context()
    bufs = Ptr{Cvoid}[]
    try
        rowptr = _upload(c, Int64.(Dt.colptr .- 1)); push!(bufs, rowptr)
        col = _upload(c, Int32.(Dt.rowval .- 1)); push!(bufs, col)
        vals = _upload(c, Vector{Float64}(Dt.nzval)); push!(bufs, vals)
        b = _upload(c, Vector{Float64}(f)); push!(bufs, b)
        x = _upload(c, zeros(N)); push!(bufs, x)
        iters = Ref{Int64}(0)
        relres = Ref{Float64}(0.0)
        check(ccall((:rbffd_lsq_solve_device, LIB), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                     Int32, Float64, Int64, Ref{Int64}, Ref{Float64}),
                    c, M, N, rowptr, col, vals, b, x, SOLVE_METHODS[method], tol, max_iters, iters, relres))
        out = Vector{Float64}(undef, N)
        GC.@preserve out check(ccall((:rbffd_memcpy_d2h, LIB), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64),
                                     c, pointer(out), x, sizeof(out)))
        return out
    finally
        foreach(p -> _dfree(c, p), bufs)
    end
end

end # module
