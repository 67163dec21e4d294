"""
    QuSpinGPU

Julia binding of libqsgpu.so (include/qsgpu.h) for QuSpin.jl.  It adds a
`GPUQuantumLinearOperator` that satisfies the same AbstractMatrix protocol as
`QuantumLinearOperator` (operators.jl:3128-3245), so ARPACK, the Krylov
exponential and BlockOps consume it unchanged, plus `eigsh` / `evolve` /
`expm_multiply` methods that keep the Krylov basis resident on the GPU.

Julia is not installed where this repository is built, so this file is
exercised only through the same C ABI by the Python tests.
"""
module QuSpinGPU

using LinearAlgebra
using SparseArrays
import QuSpin: SpinBasis1D, OperatorTerm, QuantumLinearOperator, eigsh, evolve

export GPUBasis, GPUBasisGeneral, GPUQuantumLinearOperator, GPUHamiltonian, GPUBlockOps, assemble!, block_expm

const libqsgpu = get(ENV, "QSGPU_LIB", joinpath(@__DIR__, "..", "_build", "libqsgpu.so"))

const QS_OK, QS_EARG, QS_EDIM, QS_ENOCONV, QS_EOOM, QS_EDEVICE = 0, 1, 2, 3, 4, 5
const QS_UNSET = Cint(-1000000)
const QS_F64, QS_C128, QS_F32, QS_C64 = Cint(0), Cint(1), Cint(2), Cint(3)
const WHICH = Dict(:SA => 0, :LA => 1, :SM => 2, :LM => 3, :BE => 4)

function last_error()
    buf = Vector{UInt8}(undef, 1024)
    ccall((:qs_last_error, libqsgpu), Csize_t, (Ptr{UInt8}, Csize_t), buf, length(buf))
    return unsafe_string(pointer(buf))
end

# status code -> the exception the reference throws for the same condition
function check(status::Integer)
    status == QS_OK && return nothing
    msg = last_error()
    status == QS_EARG && throw(ArgumentError(msg))
    status == QS_EDIM && throw(DimensionMismatch(msg))
    status == QS_ENOCONV && throw(ErrorException(msg))
    throw(ErrorException("libqsgpu status $status: $msg"))
end

_blk(v) = v === nothing ? QS_UNSET : Cint(v)
_dtype(::Type{Float64}) = QS_F64
_dtype(::Type{ComplexF64}) = QS_C128
_dtype(::Type{Float32}) = QS_F32
_dtype(::Type{ComplexF32}) = QS_C64

"""GPU-resident sector of a SpinBasis1D (same keywords, basis.jl:622-639)."""
mutable struct GPUBasis
    handle::Ptr{Cvoid}
    L::Int
    Ns::Int
end

function GPUBasis(L::Integer; nup=nothing, pauli::Bool=true, a::Integer=1, kblock=nothing,
                  pblock=nothing, zblock=nothing, pzblock=nothing, zAblock=nothing, zBblock=nothing)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:qs_basis_create, libqsgpu), Cint,
                (Cint, Cint, Cint, Cint, Cint, Cint, Cint, Cint, Cint, Cint, Cint, Ref{Ptr{Cvoid}}),
                L, _blk(nup), pauli, a, _blk(kblock), _blk(pblock), _blk(zblock), _blk(pzblock),
                _blk(zAblock), _blk(zBblock), 0, h))
    n = Ref{Int64}(0)
    check(ccall((:qs_basis_size, libqsgpu), Cint, (Ptr{Cvoid}, Ref{Int64}), h[], n))
    b = GPUBasis(h[], Int(L), Int(n[]))
    finalizer(x -> ccall((:qs_basis_destroy, libqsgpu), Cint, (Ptr{Cvoid},), x.handle), b)
    return b
end

"""
    GPUBasis(basis::SpinBasis1D)

The GPU sector of a reference basis, from the keywords the reference stores:
`L`, `nup` and `pauli` (basis.jl:206-213) and the symmetry blocks it records
in `basis.symmetry.blocks` (`:a`, `:kblock`, `:pblock`, `:zblock`, `:pzblock`,
`:zAblock`, `:zBblock`; basis.jl:671-760).  The GPU enumerates the same
sector, so its dimension must equal `length(basis)`; a mismatch raises.
"""
function GPUBasis(basis::SpinBasis1D)
    blk = basis.symmetry.blocks
    g(name) = get(blk, name, nothing)
    b = GPUBasis(getfield(basis, :L); nup=getfield(basis, :nup), pauli=getfield(basis, :pauli),
                 a=something(g(:a), 1), kblock=g(:kblock), pblock=g(:pblock), zblock=g(:zblock),
                 pzblock=g(:pzblock), zAblock=g(:zAblock), zBblock=g(:zBblock))
    b.Ns == length(basis) ||
        throw(DimensionMismatch("GPU sector has $(b.Ns) states, the reference basis $(length(basis))"))
    return b
end

"""
    GPUBasisGeneral(L; nup=nothing, pauli=true, block_order=nothing, blocks...)

The GPU sector of `SpinBasisGeneral(L; nup, pauli, block_order, blocks...)`
(basis.jl:530-620, spin 1/2): each block `name=(map, quantum_number)` with the
map in QuSpin's zero-based or Julia's one-based site labels, stages in
`block_order` (default: keyword order).  The reference basis keeps only the
quantum numbers, so the maps are given again here.  Up to 256 sites: beyond
32 sites with blocks, or beyond 63 (the reference's WideSpinBasis), a wide
basis that needs `nup`; its states beyond 64 sites come back as UInt256-sized
`BigInt`s from `states`.
"""
function GPUBasisGeneral(L::Integer; nup=nothing, pauli::Bool=true, block_order=nothing, blocks...)
    names = block_order === nothing ? Symbol[Symbol(n) for (n, _) in blocks] : Symbol.(collect(block_order))
    spec = Dict{Symbol,Any}(Symbol(n) => v for (n, v) in blocks)
    for n in keys(spec)
        n in names || push!(names, n)
    end
    all(n -> haskey(spec, n), names) || throw(ArgumentError("block_order refers to an unknown symmetry block"))
    maps = Matrix{Int32}(undef, L, length(names))
    qn = Vector{Int32}(undef, length(names))
    for (j, n) in enumerate(names)
        m, q = spec[n][1], spec[n][2]
        length(m) == L || throw(ArgumentError("a general symmetry map must contain one entry per site"))
        maps[:, j] .= Int32.(collect(m))
        qn[j] = Int32(q)
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:qs_basis_create_general, libqsgpu), Cint,
                (Cint, Cint, Cint, Cint, Ptr{Int32}, Ptr{Int32}, Cint, Ref{Ptr{Cvoid}}),
                L, _blk(nup), pauli, length(names), maps, qn, 0, h))
    n = Ref{Int64}(0)
    check(ccall((:qs_basis_size, libqsgpu), Cint, (Ptr{Cvoid}, Ref{Int64}), h[], n))
    b = GPUBasis(h[], Int(L), Int(n[]))
    finalizer(x -> ccall((:qs_basis_destroy, libqsgpu), Cint, (Ptr{Cvoid},), x.handle), b)
    return b
end

function states(b::GPUBasis)
    if b.L > 64
        # four 64-bit words per state (site i = bit i % 64 of word i / 64)
        w = Vector{UInt64}(undef, 4 * b.Ns)
        check(ccall((:qs_basis_states_wide, libqsgpu), Cint, (Ptr{Cvoid}, Cint, Ptr{UInt64}), b.handle, 0, w))
        return [sum(BigInt(w[4 * (i - 1) + k + 1]) << (64 * k) for k in 0:3) for i in 1:b.Ns]
    end
    out = Vector{UInt64}(undef, b.Ns)
    check(ccall((:qs_basis_states, libqsgpu), Cint, (Ptr{Cvoid}, Ptr{UInt64}), b.handle, out))
    return out
end

"""
    GPUQuantumLinearOperator(basis, static_list; diagonal, check_symm, check_herm, check_pcon)

Drop-in for QuantumLinearOperator (operators.jl:2492-2552) on the GPU.
"""
mutable struct GPUQuantumLinearOperator{T<:Number} <: AbstractMatrix{T}
    handle::Ptr{Cvoid}
    basis::GPUBasis
    hermitian::Bool
    transposed::Bool
    conjugated::Bool
end

function GPUQuantumLinearOperator(basis::GPUBasis, terms::AbstractVector{<:OperatorTerm};
                                  diagonal=nothing, check_symm::Bool=true, check_herm::Bool=true,
                                  check_pcon::Bool=true)
    ops = [t.op for t in terms]
    ncoup = Int32[length(t.couplings) for t in terms]
    coef = Float64[]
    sites = Int32[]
    iscomplex = false
    for t in terms, c in t.couplings
        z = ComplexF64(first(c))
        push!(coef, real(z), imag(z))
        iscomplex |= first(c) isa Complex
        append!(sites, Int32.(Base.tail(c)))
    end
    d = diagonal === nothing ? C_NULL : reinterpret(Float64, ComplexF64.(diagonal))
    flags = (check_symm ? 1 : 0) | (check_herm ? 2 : 0) | (check_pcon ? 4 : 0) | (iscomplex ? 8 : 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:qs_op_create, libqsgpu), Cint,
                (Ptr{Cvoid}, Cint, Ptr{Cstring}, Ptr{Int32}, Ptr{Float64}, Ptr{Int32}, Ptr{Float64}, Cint,
                 Ref{Ptr{Cvoid}}),
                basis.handle, length(terms), ops, ncoup, coef, sites, d, flags, h))
    c, herm, n = Ref{Cint}(0), Ref{Cint}(0), Ref{Int64}(0)
    check(ccall((:qs_op_info, libqsgpu), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}, Ref{Int64}),
                h[], c, herm, n))
    T = c[] != 0 ? ComplexF64 : Float64
    op = GPUQuantumLinearOperator{T}(h[], basis, herm[] != 0, false, false)
    finalizer(x -> ccall((:qs_op_destroy, libqsgpu), Cint, (Ptr{Cvoid},), x.handle), op)
    return op
end

# the reference's constructor signature, QuantumLinearOperator(basis, static_list; ...)
# (operators.jl:2492-2552), on its own SpinBasis1D: unchanged host code
# replaces `QuantumLinearOperator` by `GPUQuantumLinearOperator`
GPUQuantumLinearOperator(basis::SpinBasis1D, terms::AbstractVector{<:OperatorTerm}; kwargs...) =
    GPUQuantumLinearOperator(GPUBasis(basis), terms; kwargs...)

# an existing reference operator: its basis, static list and user diagonal
# (fields of QuantumLinearOperator, operators.jl:2331-2345); the reference
# already checked the list, so the checks are not repeated
GPUQuantumLinearOperator(op::QuantumLinearOperator) =
    GPUQuantumLinearOperator(GPUBasis(op.basis), op.static_list;
                             diagonal=any(!iszero, op.diagonal) ? op.diagonal : nothing,
                             check_symm=false, check_herm=false, check_pcon=false)

"""
    GPUHamiltonian(basis, terms; static_fmt=:csr, check_symm, check_herm, check_pcon)

Hamiltonian(basis, terms; static_fmt) (operators.jl:578-614) on the GPU: the
sector matrix is assembled on the device once (`:csr` / `:csc`; the product of
a Hermitian sector matrix reads the same stored rows either way), and every
product then streams the stored rows.  `:dense` is not kept on the device:
the assembled rows serve it, and `SparseMatrixCSC(H)` / `Matrix(H)` return the
entries on the host.
"""
function GPUHamiltonian(basis::Union{SpinBasis1D,GPUBasis}, terms::AbstractVector{<:OperatorTerm};
                        static_fmt=:csr, kwargs...)
    fmt = Symbol(lowercase(String(static_fmt)))
    fmt in (:csr, :csc, :dense) || throw(ArgumentError("static_fmt must be :dense, :csr, or :csc"))
    return assemble!(GPUQuantumLinearOperator(basis, terms; kwargs...))
end

# AbstractMatrix protocol (operators.jl:3228-3245)
Base.size(op::GPUQuantumLinearOperator) = (op.basis.Ns, op.basis.Ns)
Base.getindex(op::GPUQuantumLinearOperator, i::Int, j::Int) =
    error("GPUQuantumLinearOperator is matrix-free; use mul!")
LinearAlgebra.ishermitian(op::GPUQuantumLinearOperator) = op.hermitian
LinearAlgebra.issymmetric(op::GPUQuantumLinearOperator{T}) where {T} = T <: Real && op.hermitian
Base.transpose(op::GPUQuantumLinearOperator{T}) where {T} =
    GPUQuantumLinearOperator{T}(op.handle, op.basis, op.hermitian, !op.transposed, op.conjugated)
Base.conj(op::GPUQuantumLinearOperator{T}) where {T} =
    GPUQuantumLinearOperator{T}(op.handle, op.basis, op.hermitian, op.transposed, !op.conjugated)
Base.adjoint(op::GPUQuantumLinearOperator{T}) where {T} =
    GPUQuantumLinearOperator{T}(op.handle, op.basis, op.hermitian, !op.transposed, !op.conjugated)

# mul!(y, op, x, alpha, beta) -- operators.jl:3147-3208 (host arrays are staged)
function LinearAlgebra.mul!(y::StridedVecOrMat{S}, op::GPUQuantumLinearOperator, x::StridedVecOrMat{S},
                            alpha::Number, beta::Number) where {S<:Union{Float64,ComplexF64,Float32,ComplexF32}}
    size(x, 1) == size(op, 2) || throw(DimensionMismatch("operator and value dimensions do not match"))
    size(y) == size(x) || throw(DimensionMismatch("output and operator dimensions do not match"))
    a, b = ComplexF64(alpha), ComplexF64(beta)
    check(ccall((:qs_op_matvec, libqsgpu), Cint,
                (Ptr{Cvoid}, Cint, Ptr{S}, Ptr{S}, Int64, Cdouble, Cdouble, Cdouble, Cdouble, Cint, Cint, Cint),
                op.handle, _dtype(S), x, y, size(x, 2), real(a), imag(a), real(b), imag(b),
                op.transposed, op.conjugated, 0))
    return y
end
LinearAlgebra.mul!(y::StridedVecOrMat, op::GPUQuantumLinearOperator, x::StridedVecOrMat) =
    mul!(y, op, x, true, false)
Base.:*(op::GPUQuantumLinearOperator{T}, x::AbstractVector{S}) where {T,S} =
    mul!(zeros(promote_type(T, S), size(op, 1)), op, Vector{promote_type(T, S)}(x))

# eigsh with the Krylov basis resident on the device (operators.jl:3367-3408)
function eigsh(op::GPUQuantumLinearOperator{T}; k::Integer=min(6, size(op, 1)), which=:SA,
               tol::Real=0.0, maxiter::Integer=300, ncv=nothing, v0=nothing,
               return_eigenvectors::Bool=true, dtype::Type=T) where {T}
    n = size(op, 1)
    1 <= k <= n || throw(ArgumentError("k must lie in 1:Ns"))
    w = Symbol(uppercase(String(which)))
    haskey(WHICH, w) || throw(ArgumentError("which must be SA, LA, SM, LM, or BE"))
    # a non-Hermitian operator takes Arpack.eigs' complex path (operators.jl:3383-3390)
    op.hermitian || dtype <: Complex || (dtype = complex(dtype))
    vals = zeros(Float64, op.hermitian ? k : 2k)
    vecs = return_eigenvectors ? zeros(dtype, n, k) : nothing
    v = v0 === nothing ? C_NULL : Vector{dtype}(v0)
    nconv, nmv = Ref{Cint}(0), Ref{Cint}(0)
    check(ccall((:qs_eigsh, libqsgpu), Cint,
                (Ptr{Cvoid}, Cint, Cint, Cint, Cdouble, Cint, Cint, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid},
                 Ref{Cint}, Ref{Cint}),
                op.handle, _dtype(dtype), k, WHICH[w], tol, maxiter, ncv === nothing ? 0 : ncv, v, vals,
                vecs === nothing ? C_NULL : vecs, nconv, nmv))
    values = op.hermitian ? vals : complex.(vals[1:2:end], vals[2:2:end])
    return return_eigenvectors ? (values, vecs) : values
end

# exp(scale * op) v for every scale: _krylov_expmv_times (operators.jl:1805-1893)
function expm_multiply(op::GPUQuantumLinearOperator, v::AbstractVector, scales::AbstractVector;
                       tol::Real=1e-12, krylov_dim::Integer=30, dtype::Type=ComplexF64)
    x = Vector{dtype}(v)
    sc = ComplexF64.(scales)
    out = zeros(dtype, length(x), length(sc))
    nmv = Ref{Cint}(0)
    check(ccall((:qs_expm_multiply, libqsgpu), Cint,
                (Ptr{Cvoid}, Cint, Ptr{dtype}, Int64, Ptr{ComplexF64}, Cdouble, Cint, Ptr{dtype}, Ref{Cint}),
                op.handle, _dtype(dtype), x, length(sc), sc, tol, krylov_dim, out, nmv))
    return out
end

# evolve for a static generator (operators.jl:2073-2192, :SE branch)
function evolve(op::GPUQuantumLinearOperator, v0::AbstractVector, t0::Real, times;
                imag_time::Bool=false, tol::Real=1e-12, krylov_dim::Integer=30)
    targets = times isa Number ? [times] : collect(times)
    scales = imag_time ? ComplexF64[-(t - t0) for t in targets] : ComplexF64[-im * (t - t0) for t in targets]
    states = expm_multiply(op, v0, scales; tol, krylov_dim)
    imag_time && foreach(s -> rmul!(s, inv(norm(s))), eachcol(states))
    return times isa Number ? states[:, 1] : states
end

# lanczos_full(A, v0, m; full_ortho, eps) -- tools.jl:2742-2809, recurrence on the device
function lanczos_full(op::GPUQuantumLinearOperator{T}, v0::AbstractVector, m::Integer;
                      full_ortho::Bool=false, eps=nothing) where {T}
    n = size(op, 1)
    1 <= m < n || throw(ArgumentError("Lanczos dimension m must satisfy 1 <= m < length(v0)"))
    S = promote_type(T, eltype(v0))
    x = Vector{S}(v0)
    alphas, betas = zeros(Float64, m), zeros(Float64, max(1, m - 1))
    E, V = zeros(Float64, m), zeros(Float64, m * m)
    QT = zeros(S, n, m)                   # column j = Lanczos vector j (row j of Q_T)
    used = Ref{Cint}(0)
    check(ccall((:qs_lanczos, libqsgpu), Cint,
                (Ptr{Cvoid}, Cint, Ptr{S}, Cint, Cint, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}, Ptr{S}, Ref{Cint}),
                op.handle, _dtype(S), x, m, full_ortho, eps === nothing ? 0.0 : eps, alphas, betas, E, V, QT,
                used))
    u = Int(used[])
    return E[1:u], reshape(V[1:u*u], u, u), permutedims(QT[:, 1:u])
end

# Hamiltonian(...; static_fmt=:csr) -- operators.jl:461-576: the sector matrix
# assembled on the device; plain products then read the stored rows
function assemble!(op::GPUQuantumLinearOperator)
    check(ccall((:qs_op_assemble, libqsgpu), Cint, (Ptr{Cvoid},), op.handle))
    return op
end

# the assembled matrix with duplicates merged and exact zeros dropped
# (sparse_formats.jl:21-50 stores the same entries row by row)
function SparseArrays.SparseMatrixCSC(op::GPUQuantumLinearOperator{T}) where {T}
    n = size(op, 1)
    nnz = Ref{Int64}(0)
    check(ccall((:qs_op_csr, libqsgpu), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                op.handle, nnz, C_NULL, C_NULL, C_NULL))
    rowptr, col, val = Vector{Int64}(undef, n + 1), Vector{Int64}(undef, nnz[]), Vector{Float64}(undef, 2nnz[])
    check(ccall((:qs_op_csr, libqsgpu), Cint, (Ptr{Cvoid}, Ref{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                op.handle, nnz, rowptr, col, val))
    z = reinterpret(ComplexF64, val)
    rows = [r for r in 1:n for _ in 1:(rowptr[r+1] - rowptr[r])]
    return sparse(rows, col .+ 1, T <: Real ? real.(z) : T.(z), n, n)
end
Base.Matrix(op::GPUQuantumLinearOperator) = Matrix(SparseMatrixCSC(op))

"""
    GPUBlockOps(blocks, static, basis_args; basis_kwargs...)

BlockOps (tools.jl:2303-2353) over GPU sectors: one `GPUBasis` and one
`GPUQuantumLinearOperator` per non-empty block.  `evolve` and `block_expm`
run every active block in one lockstep batch (`qs_blocks_evolve`).
"""
struct GPUBlockOps
    bases::Vector{GPUBasis}
    ops::Vector{GPUQuantumLinearOperator}
    kinds::Vector{Cint}
end

function GPUBlockOps(blocks, static::AbstractVector{<:OperatorTerm}, basis_args::Tuple; basis_kwargs...)
    bases, ops, kinds = GPUBasis[], GPUQuantumLinearOperator[], Cint[]
    parent_nup = get(basis_kwargs, :nup, nothing)
    for block in blocks
        b = try
            GPUBasis(basis_args...; basis_kwargs..., block...)
        catch e
            e isa ArgumentError && occursin("empty", e.msg) && continue
            rethrow()
        end
        push!(bases, b)
        push!(ops, GPUQuantumLinearOperator(b, static))
        # blocks over nup inside a parent without one: the parent vector is indexed by the state
        push!(kinds, (haskey(block, :nup) && parent_nup === nothing) ? Cint(1) : Cint(0))
    end
    return GPUBlockOps(bases, ops, kinds)
end

function _blocks_evolve(bo::GPUBlockOps, psi::AbstractVector, scales::AbstractVector; tol=1e-12, krylov_dim=30)
    x = Vector{ComplexF64}(psi)
    sc = ComplexF64.(scales)
    out = zeros(ComplexF64, length(x), length(sc))
    check(ccall((:qs_blocks_evolve, libqsgpu), Cint,
                (Ptr{Ptr{Cvoid}}, Ptr{Cint}, Cint, Cint, Ptr{ComplexF64}, Int64, Ptr{ComplexF64}, Cdouble, Cint,
                 Ptr{ComplexF64}, Ptr{Int32}),
                [o.handle for o in bo.ops], bo.kinds, length(bo.ops), QS_C128, x, length(sc), sc, tol,
                krylov_dim, out, C_NULL))
    return out
end

# evolve(::BlockOps, psi_0, t0, times) -- tools.jl:2588-2646 (static generator)
function evolve(bo::GPUBlockOps, psi_0::AbstractVector, t0::Real, times; imag_time::Bool=false,
                tol::Real=1e-12, krylov_dim::Integer=30)
    targets = times isa Number ? [times] : collect(times)
    scales = imag_time ? ComplexF64[-(t - t0) for t in targets] : ComplexF64[-im * (t - t0) for t in targets]
    out = _blocks_evolve(bo, psi_0, scales; tol, krylov_dim)
    return times isa Number ? out[:, 1] : out
end

# block_expm(::BlockOps, psi_0; a, start, stop, num, endpoint) -- tools.jl:2656-2702
function block_expm(bo::GPUBlockOps, psi_0::AbstractVector; a::Number=-im, start=nothing, stop=nothing,
                    num::Integer=50, endpoint::Bool=true)
    grid = start === nothing && stop === nothing ? [1.0] :
        collect(range(start, stop; length=num + (!endpoint)))
    !endpoint && pop!(grid)
    out = _blocks_evolve(bo, psi_0, ComplexF64.(grid .* a))
    return length(grid) == 1 ? vec(out) : out
end

end # module
