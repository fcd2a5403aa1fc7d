"""
    SyntheticscRNAseqHIP

Julia drop-in for the GPU CLE path of SyntheticscRNAseq.jl on AMD MI355X
(gfx950).  It adds the same methods the CUDA package extension adds
(ext/SyntheticscRNAseqCUDAExt.jl:48-52, 132-136, 409-412) and forwards
them to `libscrna.so` through `ccall`; all simulation runs in that
library's HIP kernels.  Load it next to the package:

    using SyntheticscRNAseq, SyntheticscRNAseqHIP
    Y = simulate(net, CLE(0.01), kin, Val(:gpu); cell_num=100_000, T=50.0)

Set `ENV["SCRNA_LIB"]` to the library path if it is not next to this file.
"""
module SyntheticscRNAseqHIP

using SyntheticscRNAseq
using SyntheticscRNAseq: GeneNetwork, KineticParams, PopulationConfig, CaptureModel, CLE, BinomialTauLeap
using Random

const LIB = get(ENV, "SCRNA_LIB", joinpath(@__DIR__, "..", "libscrna.so"))

# ---- C structs (include/scrna.h) --------------------------------------------
struct CNetwork
    G::Int32
    basals::Ptr{Float64}
    A::Ptr{Float64}
    n_coop::Int32
    coop_target::Ptr{Int32}
    coop_off::Ptr{Int32}
    coop_src::Ptr{Int32}
    coop_strength::Ptr{Float64}
    n_redun::Int32
    redun_target::Ptr{Int32}
    redun_off::Ptr{Int32}
    redun_src::Ptr{Int32}
    redun_strength::Ptr{Float64}
    k_on::Ptr{Float64}
    k_off::Ptr{Float64}
end

struct CKinetics
    k_t::Float64; K_d::Float64; n::Float64; mu_m::Float64; mu_p::Float64; dilution::Float64
end

struct CPopulation
    enabled::Int32; growth_rate::Float64; V_div::Float64; V_init_lo::Float64; V_init_hi::Float64
    div_check_interval::Int32; partition::Int32; replace_mode::Int32; growth_noise::Float64
end

struct CCapture
    enabled::Int32; efficiency::Float64; efficiency_std::Float64; readout::Int32
end

struct CRun
    dt::Float64; T::Float64; cell_num::Int64; readout::Int32; init_mode::Int32; seed::UInt64
    net_id::UInt32; precision::Int32; snapshot_every::Int64; injected_noise::Ptr{Float32}
    algorithm::Int32; reserved0::Int32
end

struct COutput
    y::Ptr{Float64}; ld::Int64; counts::Ptr{Int32}; snapshots::Ptr{Float32}; moments::Ptr{Float64}
    volumes::Ptr{Float64}
end

const READOUT = Dict(:mrna => Int32(0), :protein => Int32(1), :both => Int32(2))
const PRECISION = Dict(:fp32 => Int32(0), :bf16 => Int32(1), :bf16x3 => Int32(2), :fp32_unfused => Int32(3),
                       :fp32_mfma => Int32(4), :fp32_valu => Int32(5))

# ---- context ------------------------------------------------------------------
const CTX = Ref{Ptr{Cvoid}}(C_NULL)

function _check(rc::Integer, what::AbstractString)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:scrna_last_error, LIB), Cstring, ()))
    rc == -1 && throw(ArgumentError("$what: $msg"))       # invalid argument
    rc == -6 && error("$what: $msg")                       # unsupported combination
    error("$what failed ($rc): $msg")
end

function context()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        _check(ccall((:scrna_ctx_create, LIB), Cint, (Int32, UInt64, Ptr{Ptr{Cvoid}}), 0, 0, h),
               "scrna_ctx_create")
        CTX[] = h[]
    end
    return CTX[]
end

# Flatten a GeneNetwork (1-based) into 0-based C arrays; `keep` roots them.
function _cnet(net::GeneNetwork, keep::Vector{Any})
    A = Matrix{Float64}(net.interactions)          # column-major, as the C ABI expects
    basals = Vector{Float64}(net.basals)
    push!(keep, A, basals)
    function gates(edges)
        isempty(edges) && return (Int32(0), Ptr{Int32}(0), Ptr{Int32}(0), Ptr{Int32}(0), Ptr{Float64}(0))
        tgt = Int32[e.target - 1 for e in edges]
        off = Int32[0; cumsum([length(e.sources) for e in edges])]
        src = Int32[s - 1 for e in edges for s in e.sources]
        st = Float64[e.strength for e in edges]
        push!(keep, tgt, off, src, st)
        return (Int32(length(edges)), pointer(tgt), pointer(off), pointer(src), pointer(st))
    end
    c = gates(net.cooperative)
    r = gates(net.redundant)
    kon = Vector{Float64}(net.k_on); koff = Vector{Float64}(net.k_off)
    push!(keep, kon, koff)
    return CNetwork(Int32(net.G), pointer(basals), pointer(A), c..., r..., pointer(kon), pointer(koff))
end

_ckin(k::KineticParams) = CKinetics(k.k_t, k.K_d, k.n, k.mu_m, k.mu_p, k.dilution)

function _cpop(pop::Union{Nothing,PopulationConfig}; partition=:halve, replace=:uniform)
    pop === nothing && return CPopulation(0, 0, 0, 0, 0, 0, 0, 0, 0)
    CPopulation(1, pop.growth_rate, pop.V_div, pop.V_init[1], pop.V_init[2], Int32(pop.div_check_interval),
                partition === :binomial ? 1 : 0, replace === :exclude_mother ? 1 : 0, pop.growth_noise)
end

_ccap(c::Union{Nothing,CaptureModel}) = c === nothing ? CCapture(0, 0, 0, 0) :
    CCapture(1, c.efficiency, c.efficiency_std, READOUT[c.readout])

const GpuAlg = Union{CLE,BinomialTauLeap}
_algorithm(::CLE) = Int32(0)
_algorithm(::BinomialTauLeap) = Int32(1)

function _crun(alg::GpuAlg, cell_num, T, readout, seed, net_id; init=:steady, precision=:fp32, snapshot_every=0)
    CRun(alg.dt, T, Int64(cell_num), READOUT[readout], init === :zeros ? 1 : 0, UInt64(seed), UInt32(net_id),
         PRECISION[precision], Int64(snapshot_every), Ptr{Float32}(0), _algorithm(alg), Int32(0))
end

# ---- drop-in methods -------------------------------------------------------------

# ext/SyntheticscRNAseqCUDAExt.jl:48-117
function SyntheticscRNAseq.simulate(net::GeneNetwork, alg::CLE, kin::KineticParams, ::Val{:gpu};
                                    cell_num::Int=1000, T::Float64=50.0, readout::Symbol=:protein,
                                    seed::UInt64=rand(Random.RandomDevice(), UInt64), init::Symbol=:steady,
                                    precision::Symbol=:fp32)
    _simulate(net, alg, kin, nothing, nothing; cell_num, T, readout, seed, init, precision)[1]
end

# ext/SyntheticscRNAseqCUDAExt.jl:230-311 (exact Poisson / binomial draws; counts as Float64)
function SyntheticscRNAseq.simulate(net::GeneNetwork, alg::BinomialTauLeap, kin::KineticParams, ::Val{:gpu};
                                    cell_num::Int=1000, T::Float64=50.0, readout::Symbol=:protein,
                                    seed::UInt64=rand(Random.RandomDevice(), UInt64), init::Symbol=:steady)
    _simulate(net, alg, kin, nothing, nothing; cell_num, T, readout, seed, init)[1]
end

# ext/SyntheticscRNAseqCUDAExt.jl:543-680 (counts, binomial partitioning, division on the device;
# the Hill functions read counts as in tauleap_binomial.jl:213-221)
function SyntheticscRNAseq.simulate(net::GeneNetwork, alg::BinomialTauLeap, kin::KineticParams, ::Val{:gpu},
                                    pop::PopulationConfig; T::Float64=50.0, readout::Symbol=:protein,
                                    seed::UInt64=rand(Random.RandomDevice(), UInt64))
    _simulate(net, alg, kin, pop, nothing; cell_num=pop.cell_num, T, readout, seed, partition=:binomial)[1]
end

# ext/SyntheticscRNAseqCUDAExt.jl:409-529 (division on the device, no host round trip)
function SyntheticscRNAseq.simulate(net::GeneNetwork, alg::CLE, kin::KineticParams, ::Val{:gpu},
                                    pop::PopulationConfig; T::Float64=50.0, readout::Symbol=:protein,
                                    seed::UInt64=rand(Random.RandomDevice(), UInt64), partition::Symbol=:halve)
    _simulate(net, alg, kin, pop, nothing; cell_num=pop.cell_num, T, readout, seed, partition)[1]
end

# North-star keyword form: simulate(net, kin; algorithm=CLE(dt), capture=CaptureModel(...)) -> counts
function SyntheticscRNAseq.simulate(net::GeneNetwork, kin::KineticParams; algorithm::GpuAlg,
                                    capture::CaptureModel, cell_num::Int=1000, T::Float64=50.0,
                                    seed::UInt64=rand(Random.RandomDevice(), UInt64), init::Symbol=:steady,
                                    precision::Symbol=:fp32, population=nothing)
    n = population === nothing ? cell_num : population.cell_num
    _simulate(net, algorithm, kin, population, capture; cell_num=n, T, readout=capture.readout, seed, init,
              precision)[2]
end

function _simulate(net, alg, kin, pop, cap; cell_num, T, readout, seed, init=:steady, precision=:fp32,
                   partition=:halve)
    keep = Any[]
    cn = Ref(_cnet(net, keep))
    ck = Ref(_ckin(kin))
    cr = Ref(_crun(alg, cell_num, T, readout, seed, 0; init, precision))
    cp = Ref(_cpop(pop; partition))
    cc = Ref(_ccap(cap))
    ncol = readout === :both ? 2 * net.G : net.G
    Y = Matrix{Float64}(undef, cell_num, ncol)
    counts = cap === nothing ? Matrix{Int32}(undef, 0, 0) : Matrix{Int32}(undef, cell_num, net.G)
    out = Ref(COutput(pointer(Y), Int64(cell_num), cap === nothing ? Ptr{Int32}(0) : pointer(counts),
                      Ptr{Float32}(0), Ptr{Float64}(0), Ptr{Float64}(0)))
    GC.@preserve keep Y counts begin
        rc = ccall((:scrna_simulate_cle, LIB), Cint,
                   (Ptr{Cvoid}, Ref{CNetwork}, Ref{CKinetics}, Ref{CRun}, Ref{CPopulation}, Ref{CCapture},
                    Ref{COutput}),
                   context(), cn, ck, cr, cp, cc, out)
    end
    _check(rc, "simulate")
    return Y, Matrix{Int}(counts)
end

# No counterpart in the reference: the CLE of one network without population
# dynamics, with every device buffer taken from the library's fixed per-device
# array (scrna_simulate_cle_noalloc), for callers that must not allocate device
# memory.  Larger runs than the array holds throw ArgumentError; calls on one
# device are serialised.
function simulate_noalloc(net::GeneNetwork, alg::CLE, kin::KineticParams; cell_num::Int=1000, T::Float64=50.0,
                          readout::Symbol=:protein, seed::UInt64=rand(Random.RandomDevice(), UInt64),
                          init::Symbol=:steady, precision::Symbol=:fp32, capture=nothing)
    keep = Any[]
    cn = Ref(_cnet(net, keep))
    ck = Ref(_ckin(kin))
    cr = Ref(_crun(alg, cell_num, T, readout, seed, 0; init, precision))
    cc = Ref(_ccap(capture))
    ncol = readout === :both ? 2 * net.G : net.G
    Y = Matrix{Float64}(undef, cell_num, ncol)
    counts = capture === nothing ? Matrix{Int32}(undef, 0, 0) : Matrix{Int32}(undef, cell_num, net.G)
    out = Ref(COutput(pointer(Y), Int64(cell_num), capture === nothing ? Ptr{Int32}(0) : pointer(counts),
                      Ptr{Float32}(0), Ptr{Float64}(0), Ptr{Float64}(0)))
    GC.@preserve keep Y counts begin
        rc = ccall((:scrna_simulate_cle_noalloc, LIB), Cint,
                   (Ptr{Cvoid}, Ref{CNetwork}, Ref{CKinetics}, Ref{CRun}, Ref{CCapture}, Ref{COutput}),
                   context(), cn, ck, cr, cc, out)
    end
    _check(rc, "simulate_noalloc")
    return capture === nothing ? Y : (Y, Matrix{Int}(counts))
end

# ext/SyntheticscRNAseqCUDAExt.jl:132-213 and 317-399 (networks may differ in G here)
function SyntheticscRNAseq.simulate_gpu_batch(networks::Vector{GeneNetwork}, alg::GpuAlg, kin::KineticParams;
                                              cell_num::Int=1000, T::Float64=50.0, readout::Symbol=:protein,
                                              seed::UInt64=rand(Random.RandomDevice(), UInt64))
    n = length(networks)
    keep = Any[]
    cns = [_cnet(net, keep) for net in networks]
    cks = fill(_ckin(kin), n)
    crs = [_crun(alg, cell_num, T, readout, seed, i - 1) for i in 1:n]
    cps = fill(_cpop(nothing), n)
    ccs = fill(_ccap(nothing), n)
    Ys = [Matrix{Float64}(undef, cell_num, readout === :both ? 2 * net.G : net.G) for net in networks]
    outs = [COutput(pointer(Y), Int64(cell_num), Ptr{Int32}(0), Ptr{Float32}(0), Ptr{Float64}(0),
                    Ptr{Float64}(0)) for Y in Ys]
    GC.@preserve keep Ys cns cks crs cps ccs outs begin
        rc = ccall((:scrna_simulate_cle_batch, LIB), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{CNetwork}, Ptr{CKinetics}, Ptr{CRun}, Ptr{CPopulation},
                    Ptr{CCapture}, Ptr{COutput}),
                   context(), Int32(n), cns, cks, crs, cps, ccs, outs)
    end
    _check(rc, "simulate_gpu_batch")
    return Ys
end

# GPU form of apply_capture (capture.jl:25-47)
function gpu_apply_capture(X::Matrix{<:Real}, capture::CaptureModel;
                           seed::UInt64=rand(Random.RandomDevice(), UInt64))
    Xd = Matrix{Float64}(X)
    cells, G = size(Xd)
    Y = Matrix{Int32}(undef, cells, G)
    GC.@preserve Xd Y begin
        rc = ccall((:scrna_apply_capture, LIB), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Int64, Int32, Int64, Float64, Float64, UInt64, UInt32, Ptr{Int32}),
                   context(), Xd, cells, Int32(G), cells, capture.efficiency, capture.efficiency_std, seed,
                   UInt32(0), Y)
    end
    _check(rc, "apply_capture")
    return Matrix{Int}(Y)
end

# ---- network prior (kinetics.jl:105-205) and theta constructor (network.jl:233-256) ----
struct CPrior
    G::Int32; regulation::Int32; n_tf_lo::Int32; n_tf_hi::Int32; targets_lo::Int32; targets_hi::Int32
    coop_lo::Int32; coop_hi::Int32; redun_lo::Int32; redun_hi::Int32; infer_basals::Int32
end

const REGULATION = Dict(:mixed => Int32(0), :additive => Int32(1), :cooperative => Int32(2), :redundant => Int32(3))
_rng(r) = r === nothing ? (Int32(-1), Int32(-1)) : (Int32(first(r)), Int32(last(r)))

"""
    sample_network_native(G; seed, index=0, regulation=:mixed, n_tf_range=nothing, ...)

sample_network's prior (kinetics.jl:105-181) drawn natively; network `index`
of `seed` is a pure function of the two (Philox stream), for SBI batches.
"""
function sample_network_native(G::Int; seed::UInt64, index::Integer=0, regulation::Symbol=:mixed,
                               n_tf_range=nothing, targets_per_tf_range=nothing, n_coop_range=nothing,
                               n_redun_range=nothing, infer_basals::Bool=true)
    pr = Ref(CPrior(Int32(G), REGULATION[regulation], _rng(n_tf_range)..., _rng(targets_per_tf_range)...,
                    _rng(n_coop_range)..., _rng(n_redun_range)..., Int32(infer_basals)))
    mg = Ref{Int32}(0)
    _check(ccall((:scrna_prior_max_gates, LIB), Cint, (Ref{CPrior}, Ref{Int32}), pr, mg), "sample_network")
    m = Int(mg[])
    basals = Vector{Float64}(undef, G); A = Matrix{Float64}(undef, G, G)
    gtype = Vector{Int32}(undef, m); gtarget = Vector{Int32}(undef, m); gsrc = Vector{Int32}(undef, 2m)
    gstr = Vector{Float64}(undef, m); ng = Ref{Int32}(0)
    rc = ccall((:scrna_sample_network, LIB), Cint,
               (Ref{CPrior}, UInt64, UInt32, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32},
                Ptr{Float64}, Int32, Ref{Int32}),
               pr, seed, UInt32(index), basals, A, gtype, gtarget, gsrc, gstr, Int32(m), ng)
    _check(rc, "sample_network")
    coop = SyntheticscRNAseq.CoopEdge[]; redun = SyntheticscRNAseq.RedunEdge[]
    for k in 1:ng[]
        t = Int(gtarget[k]) + 1; src = [Int(gsrc[2k - 1]) + 1, Int(gsrc[2k]) + 1]
        gtype[k] == 0 ? push!(coop, SyntheticscRNAseq.CoopEdge(t, src, gstr[k])) :
                        push!(redun, SyntheticscRNAseq.RedunEdge(t, src, gstr[k]))
    end
    return GeneNetwork(G, basals, A, coop, redun, fill(Inf, G), fill(Inf, G))
end

"""GeneNetwork(theta, G) (network.jl:233-256), natively."""
function network_from_theta_native(theta::Vector{Float64}, G::Int)
    basals = Vector{Float64}(undef, G); A = Matrix{Float64}(undef, G, G)
    _check(ccall((:scrna_network_from_theta, LIB), Cint, (Int32, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}),
                 Int32(G), theta, Int64(length(theta)), basals, A), "network_from_theta")
    return GeneNetwork(G, basals, A)
end

"""
    write_10x(dir, counts; gene_names=nothing, barcode_prefix=nothing) -> nnz

Write a cells x G count matrix as a 10x Genomics v2 directory (matrix.mtx,
genes.tsv, barcodes.tsv), the layout experiments/compare_real_scrnaseq.jl:62-96 reads.
"""
function write_10x(dir::AbstractString, counts::AbstractMatrix{<:Integer}; gene_names=nothing,
                   barcode_prefix=nothing)
    Y = Matrix{Int32}(counts)
    cells, G = size(Y)
    names = gene_names === nothing ? C_NULL : [Base.unsafe_convert(Cstring, n) for n in gene_names]
    nnz = Ref{Int64}(0)
    GC.@preserve gene_names begin
        rc = ccall((:scrna_write_10x, LIB), Cint,
                   (Cstring, Ptr{Int32}, Int64, Int32, Int64, Ptr{Cstring}, Cstring, Ref{Int64}),
                   dir, Y, Int64(cells), Int32(G), Int64(cells), names,
                   barcode_prefix === nothing ? C_NULL : barcode_prefix, nnz)
    end
    _check(rc, "write_10x")
    return nnz[]
end

function __init__()
    isfile(LIB) || @warn "libscrna.so not found at $LIB; set ENV[\"SCRNA_LIB\"]"
end

end # module
