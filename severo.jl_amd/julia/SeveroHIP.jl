# SeveroHIP.jl -- Julia binding of libsevero_hip.so (MI355X / gfx950) for
# Severo.jl's embedding-and-graph path.
#
# Two ways to use it:
#   1. Drop-in at the native boundary: point Severo's `libcell` at this
#      library (it exports irlba and FindNeighboursEuclidean32/64 and
#      FindNeighboursCosine32/64 with libcell's signatures,
#      src/irlba.jl:66-71, src/neighbours.jl:32-72).
#   2. The functions below, which take and return the same types as
#      Severo's API (normalize_cells, scale_features, embedding/pca,
#      nearest_neighbours, jaccard_index, shared_nearest_neighbours,
#      cluster) and run every stage on the GPU.  (The Python package's
#      `Pipeline` keeps a dataset resident on the device across stages.)
module SeveroHIP

using SparseArrays, LinearAlgebra
import NamedArrays: NamedArray, NamedMatrix, names, dimnames
import Severo

const libsev = get(ENV, "SEVERO_HIP_LIB", joinpath(@__DIR__, "..", "libsevero_hip.so"))

const SEV_FLOAT32, SEV_FLOAT64 = Cint(0), Cint(1)
const SEV_INT32, SEV_INT64 = Cint(0), Cint(1)
const SEV_PREC = Dict(:fp32 => Cint(0), :bf16 => Cint(1), :exact => Cint(2), :fp8 => Cint(3))

function check(rc::Integer)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sev_last_error, libsev), Cstring, ()))
    rc == -3 && error("convergence failed: $msg")
    error("libsevero_hip: $msg (status $rc)")
end

dtcode(::Type{Float32}) = SEV_FLOAT32
dtcode(::Type{Float64}) = SEV_FLOAT64

# ---------------------------------------------------------------- normalize
# src/normalize.jl:40-55
function normalize_cells(X::SparseMatrixCSC{<:Integer,Int64}; method=:lognormalize,
                         scale_factor::Real=1.0, dtype::Type{T}=Float64) where {T<:AbstractFloat}
    method = Symbol(method)
    meth = method == :lognormalize ? Cint(0) :
           method == :relativecounts ? Cint(1) : error("unknown normalization method: $method")
    m, n = size(X)
    nz = eltype(X) == Int32 ? nonzeros(X) : convert(Vector{Int64}, nonzeros(X))
    ct = eltype(nz) == Int32 ? SEV_INT32 : SEV_INT64
    out = Vector{T}(undef, nnz(X))
    check(ccall((:sev_normalize_cells, libsev), Cint,
                (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Cint, Cint, Cdouble, Cint,
                 Ptr{Cvoid}),
                m, n, X.colptr, X.rowval, nz, ct, meth, scale_factor, dtcode(T), out))
    SparseMatrixCSC(m, n, copy(X.colptr), copy(X.rowval), out)
end

normalize_cells(X::NamedArray; kw...) = NamedArray(normalize_cells(X.array; kw...), X.dicts,
                                                   X.dimnames)

# ------------------------------------------------------ variable features
# src/variablefeatures.jl:128-161: method :vst (the loess is the exact local
# fit at every gene), :dispersion, :meanvarplot (num_bins, width binning) or
# :saunders (alpha_thresh, 0.1 as the reference's keyword default)
const SEV_HVG = Dict(:dispersion => Cint(1), :meanvarplot => Cint(2), :saunders => Cint(3))

function find_variable_features(counts::NamedArray, nfeatures::Integer=2000; method=:vst,
                                loess_span::Real=0.5, num_bins::Integer=20,
                                binning_method=:width, alpha_thresh::Real=0.1)
    method = Symbol(method)
    X = counts.array
    m, n = size(X)
    nz = convert(Vector{Int32}, nonzeros(X))
    selected = Vector{Int64}(undef, nfeatures)
    if method == :vst
        check(ccall((:sev_find_variable_features, libsev), Cint,
                    (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int32}, Int64, Cdouble, Ptr{Int64},
                     Ptr{Cdouble}),
                    m, n, X.colptr, X.rowval, nz, nfeatures, loess_span, selected, C_NULL))
    elseif haskey(SEV_HVG, method)
        Symbol(binning_method) == :width || error("only width binning is provided")
        check(ccall((:sev_find_variable_features_method, libsev), Cint,
                    (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Int32}, Cint, Int64, Int64, Cint,
                     Cdouble, Ptr{Int64}, Ptr{Cdouble}),
                    m, n, X.colptr, X.rowval, nz, SEV_HVG[method], nfeatures, num_bins, Cint(0),
                    alpha_thresh, selected, C_NULL))
    else
        error("unknown selection method: $method")
    end
    NamedArray(selected, (names(counts, 2)[selected],), (dimnames(counts, 2),))
end

# ----------------------------------------------------------------- scaling
# src/scaling.jl:199-217, :335-357
function scale_features(X::NamedArray{T,2,SparseMatrixCSC{T,Int64}}; scale_max::Real=Inf,
                        dtype::Type{R}=T, features=nothing) where {T<:AbstractFloat,R<:AbstractFloat}
    if features !== nothing
        X = X[:, features]
    end
    A = X.array
    m, n = size(A)
    vals = convert(Vector{R}, nonzeros(A))
    out = similar(vals)
    mu = Vector{R}(undef, n)
    check(ccall((:sev_scale_features, libsev), Cint,
                (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Cint, Cdouble, Ptr{Cvoid},
                 Ptr{Cvoid}),
                m, n, A.colptr, A.rowval, vals, dtcode(R), scale_max, out, mu))
    B = SparseMatrixCSC(m, n, copy(A.colptr), copy(A.rowval), out)
    Severo.CenteredMatrix(NamedArray(B, X.dicts, X.dimnames),
                          NamedArray(mu, (X.dicts[2],), (X.dimnames[2],)))
end

# ------------------------------------------------ CenteredMatrix products
# The matrix forms of mul! (src/scaling.jl:259-294) on the GPU:
# mul!(Y, S, X, α, β), mul!(Y, S', X, α, β) and mul!(C, Sl', Sr, α, β).
# Named functions rather than mul! methods on Severo's types.
function centered_mul!(Y::StridedMatrix{Float64}, S::Severo.CenteredMatrix,
                       X::StridedMatrix{Float64}, α::Real=1.0, β::Real=0.0; trans::Bool=false)
    A, mu = Severo._A_mu(S)
    m, n = size(A)
    T = eltype(nonzeros(A))
    stride(X, 1) == 1 && stride(Y, 1) == 1 || throw(ArgumentError("column-major X and Y"))
    size(X, 1) == (trans ? m : n) && size(Y) == ((trans ? n : m), size(X, 2)) ||
        throw(DimensionMismatch())
    check(ccall((:sev_centered_mul_mat, libsev), Cint,
                (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cchar, Int64,
                 Ptr{Float64}, Int64, Ptr{Float64}, Int64, Cdouble, Cdouble),
                m, n, A.colptr, A.rowval, nonzeros(A), convert(Vector{T}, mu), dtcode(T),
                trans ? Cchar('T') : Cchar('N'), size(X, 2), X, max(stride(X, 2), 1), Y,
                max(stride(Y, 2), 1), α, β))
    Y
end

function centered_gram!(C::StridedMatrix{Float64}, Sl::Severo.CenteredMatrix,
                        Sr::Severo.CenteredMatrix, α::Real=1.0, β::Real=0.0)
    Al, mul = Severo._A_mu(Sl)
    Ar, mur = Severo._A_mu(Sr)
    size(Al, 1) == size(Ar, 1) && size(C) == (size(Al, 2), size(Ar, 2)) ||
        throw(DimensionMismatch())
    T = promote_type(eltype(nonzeros(Al)), eltype(nonzeros(Ar)))
    same = Sl === Sr
    check(ccall((:sev_centered_gram, libsev), Cint,
                (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Int64},
                 Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Float64}, Int64, Cdouble, Cdouble),
                size(Al, 1), size(Al, 2), Al.colptr, Al.rowval, convert(Vector{T}, nonzeros(Al)),
                convert(Vector{T}, mul), size(Ar, 2), same ? C_NULL : pointer(Ar.colptr),
                same ? C_NULL : pointer(Ar.rowval),
                same ? C_NULL : pointer(convert(Vector{T}, nonzeros(Ar))),
                same ? C_NULL : pointer(convert(Vector{T}, mur)), dtcode(T), C,
                max(stride(C, 2), 1), α, β))
    C
end

# -------------------------------------------------------------------- PCA
# src/irlba.jl:47-85 and src/embedding.jl:46-94 (algorithm = :irlba)
function irlba(C::Severo.CenteredMatrix, nu::Integer; tol=1e-5, maxit=1000, init=nothing,
               precision=:fp32, work=nu + 7)
    A, mu = Severo._A_mu(C)
    m, n = size(A)
    init = init === nothing ? randn(n) : convert(Vector{Float64}, init)
    T = eltype(nonzeros(A))
    vals = nonzeros(A)
    muv = convert(Vector{T}, mu)
    s = zeros(nu)
    U = zeros(m, nu)
    V = zeros(n, nu)
    iters = Ref{Int64}(0)
    nmult = Ref{Int64}(0)
    check(ccall((:sev_irlba_centered_csc, libsev), Cint,
                (Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Int64, Int64,
                 Int64, Cdouble, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint,
                 Ref{Int64}, Ref{Int64}),
                m, n, A.colptr, A.rowval, vals, muv, dtcode(T), nu, work, maxit, tol, init, s, U,
                V, SEV_PREC[precision], iters, nmult))
    SVD(U, s, V')
end

# _pca's algorithm rule (src/embedding.jl:46-53): :arpack and :irlba both run
# the device IRLBA; more than half the possible components take the dense
# route, the full Golub-Kahan basis of min(m, n) <= 128 Lanczos vectors
function pca(X::Severo.CenteredMatrix, npcs::Integer; algorithm=:arpack, kw...)
    m, n = size(X)
    npcs = min(min(m, n), npcs)
    if (npcs > 0.5 * min(m, n)) && (algorithm == :arpack || algorithm == :irlba)
        @warn "Computing too large a percentage of principal components, using standard svd instead"
        algorithm = :svd
    end
    algorithm == :tssvd && error("algorithm=:tssvd is not provided on the device")
    S = if algorithm == :svd
        min(m, n) <= 128 || error("the dense SVD route needs min(m, n) <= 128")
        irlba(X, npcs; work=min(m, n), kw...)
    elseif algorithm == :arpack || algorithm == :irlba
        irlba(X, npcs; kw...)
    else
        error("unknown algorithm: $algorithm")
    end
    Z = S.U[:, 1:npcs] * Diagonal(S.S[1:npcs])
    stdev = S.S[1:npcs] ./ sqrt(max(1, m - 1))
    latent = map(x -> string("PC-", x), 1:npcs)
    rown, coln = names(X.A)
    rowdim, _ = dimnames(X.A)
    Severo.LinearEmbedding(X, NamedArray(Z, (rown, latent), (rowdim, :latent)),
                           NamedArray(stdev, (latent,), (:latent,)),
                           NamedArray(Matrix(S.V[:, 1:npcs]), (coln, latent), (rowdim, :latent)))
end

embedding(X, n::Integer=50; method=:pca, kw...) =
    Symbol(method) == :pca ? pca(X, n; kw...) : error("unknown reduction method: $method")

# ------------------------------------------------------------- neighbours
# exact kNN behind ann! (src/neighbours.jl:19-72); indices 1-based, n x k;
# metric :euclidean or Euclidean(), :cosine or CosineDist() (unit rows,
# distance D / 2 of the unit rows)
_metric(m::Symbol) = m
_metric(m::AbstractString) = Symbol(m)
_metric(::Severo.Euclidean) = :euclidean
_metric(::Severo.CosineDist) = :cosine

function ann(X::StridedMatrix{T}, k::Integer; include_self::Bool=true,
             precision=:fp32, metric=:euclidean) where {T<:Union{Float32,Float64}}
    n, d = size(X)
    @assert stride(X, 1) == 1
    idx = Matrix{Int32}(undef, n, k)
    dist = Matrix{T}(undef, n, k)
    nfb = Ref{Int64}(0)
    m = _metric(metric)
    if m == :euclidean
        check(ccall((:sev_knn, libsev), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Int64, Cint, Int64, Cint, Cint, Ptr{Int32},
                     Ptr{Cvoid}, Ref{Int64}),
                    X, n, d, stride(X, 2), dtcode(T), k, include_self, SEV_PREC[precision], idx,
                    dist, nfb))
    elseif m == :cosine
        check(ccall((:sev_knn_cosine, libsev), Cint,
                    (Ptr{Cvoid}, Int64, Int64, Int64, Cint, Int64, Cint, Cint, Ptr{Int32},
                     Ptr{Cvoid}, Ref{Int64}),
                    X, n, d, stride(X, 2), dtcode(T), k, include_self, SEV_PREC[precision], idx,
                    dist, nfb))
    else
        error("unknown metric: $metric")
    end
    idx, dist
end

# src/neighbours.jl:76-86
function nearest_neighbours(X::NamedMatrix, k::Integer; dims=:, include_self::Bool=true, kw...)
    Xd = dims === (:) ? X.array : view(X.array, :, dims)
    idx, _ = ann(Xd, k; include_self=include_self, kw...)
    n = size(idx, 1)
    nn = sparse(vec(idx'), repeat(1:n, inner=k), trues(length(idx)))
    NamedArray(nn, (X.dicts[1], X.dicts[1]), (X.dimnames[1], X.dimnames[1]))
end
nearest_neighbours(em::Severo.LinearEmbedding, k::Integer; kw...) =
    nearest_neighbours(em.coordinates, k; kw...)

function _snn(idx::Matrix{Int32}, k::Integer, prune::Real, ::Type{T}) where {T<:AbstractFloat}
    n = size(idx, 1)
    nz = Ref{Int64}(0)
    check(ccall((:sev_snn_jaccard, libsev), Cint,
                (Int64, Int64, Ptr{Int32}, Cdouble, Cint, Ref{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Cvoid}),
                n, k, idx, prune, dtcode(T), nz, C_NULL, C_NULL, C_NULL))
    colptr = Vector{Int64}(undef, n + 1)
    rowval = Vector{Int64}(undef, nz[])
    nzval = Vector{T}(undef, nz[])
    check(ccall((:sev_snn_jaccard, libsev), Cint,
                (Int64, Int64, Ptr{Int32}, Cdouble, Cint, Ref{Int64}, Ptr{Int64}, Ptr{Int64},
                 Ptr{Cvoid}),
                n, k, idx, prune, dtcode(T), nz, colptr, rowval, nzval))
    SparseMatrixCSC(n, n, colptr, rowval, nzval)
end

# src/neighbours.jl:127-131 (k taken from the column counts)
function jaccard_index(nn::NamedArray; prune::Real=1/15, dtype::Type{T}=Float64) where {T}
    A = nn.array
    counts = diff(A.colptr)
    k = counts[1]
    all(==(k), counts) || error("every column of nn must hold k neighbours")
    idx = Matrix{Int32}(reshape(rowvals(A), k, :)')
    S = _snn(idx, k, prune, T)
    NamedArray(S, (nn.dicts[1], nn.dicts[1]), (nn.dimnames[1], nn.dimnames[1]))
end

# src/neighbours.jl:272-278, :302
function shared_nearest_neighbours(X::NamedMatrix{T}, k::Integer; dims=:, include_self::Bool=true,
                                   prune::Real=1/15, kw...) where {T}
    Xd = dims === (:) ? X.array : view(X.array, :, dims)
    idx, _ = ann(Xd, k; include_self=include_self, kw...)
    S = _snn(idx, k, prune, T)
    NamedArray(S, (X.dicts[1], X.dicts[1]), (X.dimnames[1], X.dimnames[1]))
end
shared_nearest_neighbours(em::Severo.LinearEmbedding, k::Integer; kw...) =
    shared_nearest_neighbours(em.coordinates, k; kw...)

# cluster(SNN; algorithm=:louvain, resolution, nrandomstarts, niterations,
# group_singletons) (src/clustering.jl:59-104): Louvain with parallel local
# moving on the device; `seed` fixes the random label orders.
function cluster(SNN::NamedArray{T,2,SparseMatrixCSC{T,Int64}}; algorithm=:louvain,
                 resolution::Real=0.8, nrandomstarts::Integer=10, niterations::Integer=10,
                 group_singletons::Bool=true, seed::Integer=0) where {T<:Union{Float32,Float64}}
    algorithm = Symbol(algorithm)
    algorithm == :louvain || error("unknown clustering algorithm: $algorithm")
    A = SNN.array
    n = Severo.checksquare(A)
    labels = Vector{Int64}(undef, n)
    q = Ref{Float64}(0.0)
    nc = Ref{Int64}(0)
    check(ccall((:sev_cluster_louvain, libsev), Cint,
                (Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Cvoid}, Cint, Cdouble, Int64, Int64, UInt64,
                 Cint, Ptr{Int64}, Ref{Float64}, Ref{Int64}),
                n, A.colptr, A.rowval, A.nzval, T == Float32 ? SEV_FLOAT32 : SEV_FLOAT64,
                resolution, nrandomstarts, niterations, UInt64(seed), group_singletons ? 1 : 0,
                labels, q, nc))
    NamedArray(labels, (SNN.dicts[1],), (:cells,))
end

# ---------------------------------------------------------------- 10X input
# src/input.jl:26-115 readMM: the text parsed by the library's host threads,
# then sparse(I, J, V, rows, cols) as the reference builds it
function readMM(fname::AbstractString)
    rows, cols, entries, isreal = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0), Ref{Int32}(0)
    check(ccall((:sev_mtx_info, libsev), Cint,
                (Cstring, Ref{Int64}, Ref{Int64}, Ref{Int64}, Ref{Int32}),
                fname, rows, cols, entries, isreal))
    I = Vector{Int64}(undef, entries[])
    J = Vector{Int64}(undef, entries[])
    if isreal[] != 0
        V = Vector{Float64}(undef, entries[])
        check(ccall((:sev_mtx_read, libsev), Cint,
                    (Cstring, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                    fname, I, J, C_NULL, V))
    else
        V = Vector{Int64}(undef, entries[])
        check(ccall((:sev_mtx_read, libsev), Cint,
                    (Cstring, Ptr{Int64}, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
                    fname, I, J, V, C_NULL))
    end
    sparse(I, J, V, rows[], cols[])
end

# src/input.jl:199-208 and :634-646: the counts matrix parsed on the host and
# transposed (duplicates summed) on the GPU, as Int32 counts
function read_10X(dirname::AbstractString; gene_column::Int64=2, unique_features::Bool=true)
    isdir(dirname) || throw(Severo.ParseError_10X("Directory $dirname does not exist"))
    is_v3 = isfile(joinpath(dirname, "features.tsv.gz"))
    feature_file = joinpath(dirname, is_v3 ? "features.tsv.gz" : "genes.tsv")
    barcodes_file = joinpath(dirname, is_v3 ? "barcodes.tsv.gz" : "barcodes.tsv")
    matrix_file = joinpath(dirname, is_v3 ? "matrix.mtx.gz" : "matrix.mtx")
    if !(isfile(feature_file) && isfile(barcodes_file) && isfile(matrix_file))
        throw(Severo.ParseError_10X("Couldn't locate all components: $feature_file, $barcodes_file, $matrix_file"))
    end
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:sev_pipeline_create, libsev), Cint, (Ref{Ptr{Cvoid}},), h))
    X = try
        m, n, nnz = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
        check(ccall((:sev_pipeline_load_mtx, libsev), Cint,
                    (Ptr{Cvoid}, Cstring, Ref{Int64}, Ref{Int64}, Ref{Int64}),
                    h[], matrix_file, m, n, nnz))
        colptr = Vector{Int64}(undef, n[] + 1)
        rowval = Vector{Int64}(undef, nnz[])
        nzval = Vector{Int32}(undef, nnz[])
        check(ccall((:sev_pipeline_get_counts, libsev), Cint,
                    (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}, Ptr{Int32}), h[], colptr, rowval, nzval))
        SparseMatrixCSC(m[], n[], colptr, rowval, nzval)
    finally
        ccall((:sev_pipeline_destroy, libsev), Cvoid, (Ptr{Cvoid},), h[])
    end
    barcodes = Severo._readlines(barcodes_file)
    features = Severo.readDelim(feature_file, header=false)[:, gene_column]
    Severo.convert_counts(X, features, barcodes, unique_features=unique_features)
end

end # module
