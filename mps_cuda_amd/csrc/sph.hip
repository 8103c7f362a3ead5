// sph.hip — DFSPH / implicit viscosity / advection kernels for gfx950.
//
// One thread per particle walks its column-major ELL neighbour lists (fluid then
// boundary), accumulating in registers in the reference's order, so a per-particle
// result equals (0 + fluid sum) + boundary sum exactly as the reference's two kernels and
// AtomicAdd produce it. Passes that the reference splits into several kernels over the
// same neighbour list are fused (density + alpha + gravity; delta + stiffness + error
// reduction; pressure force + velocity update; ST + XSPH + viscosity + final A.p).
//
// Quantities that stay fixed while a solver iterates are computed once per step or per
// solve and kept per neighbour slot: the kernel gradient factor G_ij (written by the alpha
// pass, read by every DF/CD iteration) and the two scalar coefficients of the implicit
// viscosity/surface-tension operator (written by the CG set-up pass), so an iteration
// streams 4-8 bytes per pair instead of re-evaluating square roots and divisions.
//
// Convergence sums: each workgroup reduces its T-rounded terms in double and writes one
// partial; a one-workgroup fold kernel adds the partials in index order (deterministic, no
// FP64 atomics) and updates a SolverDev record on the device. Kernels of later iterations
// read SolverDev::done first and exit at once, so the host can enqueue several iterations
// before it looks at the result.
#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mps {
namespace dev {

__device__ __forceinline__ uint32_t nbr_ff(const Nbr& nb, uint32_t k, uint32_t i, uint32_t N) {
    return nb.ff[size_t(k) * N + i];
}
__device__ __forceinline__ uint32_t nbr_fb(const Nbr& nb, uint32_t k, uint32_t i, uint32_t N) {
    return nb.fb[size_t(k) * N + i];
}

// Walks the wall list of slot i in list order, kWallBatch entries at a time: the batch's
// indices are loaded together, then gather(j, slot) fetches each entry's wall record, and only
// then does use(k, j, slot, record) consume them in order. A batch costs about one dependent
// memory latency instead of two per entry.
constexpr uint32_t kWallBatch = 4;
template <class Gather, class Use>
__device__ __forceinline__ void walk_walls(const Nbr& nb, uint32_t i, uint32_t N, uint32_t cb, Gather&& gather, Use&& use) {
    using Rec = decltype(gather(0u, size_t(0)));
    for (uint32_t k0 = 0; k0 < cb; k0 += kWallBatch) {
        uint32_t j[kWallBatch];
#pragma unroll
        for (uint32_t q = 0; q < kWallBatch; q++) j[q] = k0 + q < cb ? nbr_fb(nb, k0 + q, i, N) : 0u;
        Rec w[kWallBatch];
#pragma unroll
        for (uint32_t q = 0; q < kWallBatch; q++)
            if (k0 + q < cb) w[q] = gather(j[q], size_t(k0 + q) * N + i);
#pragma unroll
        for (uint32_t q = 0; q < kWallBatch; q++)
            if (k0 + q < cb) use(k0 + q, j[q], size_t(k0 + q) * N + i, w[q]);
    }
}

// Block-reduce `v` (double) in a fixed order and publish the workgroup's partial.
// COH: device-coherent store (agent scope), for the persistent solver's barrier without fences.
template <bool COH = false>
__device__ __forceinline__ void block_partial(double v, double* __restrict__ partials, uint32_t blk) {
    __shared__ double lds[kBlock / 64];
    const double bs = block_sum_d(v, lds);
    if (threadIdx.x == 0) {
        if constexpr (COH) __hip_atomic_store(partials + blk, bs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else partials[blk] = bs;
    }
}
// Loads and stores of the CG vectors that other workgroups read within one launch of the
// persistent solver: device-coherent (agent scope) when COH, so that no cache write-back or
// invalidation is needed at its barriers; plain otherwise.
template <bool COH>
__device__ __forceinline__ f3 ld3c(const f3* p) {
    if constexpr (COH) {
        const float* q = reinterpret_cast<const float*>(p);
        return {__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                __hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
    } else {
        return *p;
    }
}
template <bool COH>
__device__ __forceinline__ void st3c(f3* p, f3 v) {
    if constexpr (COH) {
        float* q = reinterpret_cast<float*>(p);
        __hip_atomic_store(q, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(q + 1, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(q + 2, v.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        *p = v;
    }
}
__device__ __forceinline__ void block_partial(double v, double* __restrict__ partials) {
    block_partial(v, partials, blockIdx.x);
}
// Sum of partials[0..n) in index-strided order, valid in thread 0 (one workgroup of
// kFoldThreads).
constexpr int kFoldThreads = 1024;
__device__ __forceinline__ double fold(const double* __restrict__ partials, uint32_t n) {
    __shared__ double lds[kFoldThreads / 64];
    double s = 0.0;
    constexpr uint32_t R = 8;  // loads in flight per thread; the adds keep the index order
    for (uint32_t k0 = threadIdx.x; k0 < n; k0 += R * blockDim.x) {
        double v[R];
#pragma unroll
        for (uint32_t q = 0; q < R; q++) {
            const uint32_t k = k0 + q * blockDim.x;
            v[q] = k < n ? partials[k] : 0.0;
        }
#pragma unroll
        for (uint32_t q = 0; q < R; q++)
            if (k0 + q * blockDim.x < n) s += v[q];
    }
    return block_sum_d(s, lds);
}

// Fluid-fluid list entries each thread keeps in flight in the halo-staged kernels (their
// occupancy is bounded by LDS, so registers are free for it). The 16-bit halo positions are
// read with ordinary loads: they are reused by every solver iteration and stay in the
// memory-side cache, while the larger per-pair coefficient streams are read non-temporally.
constexpr uint32_t kListBatch = 4;       // once-per-step walks (density/alpha, ST factor, CG setup)
constexpr uint32_t kListBatchSolve = 2;  // solver iterations (DF/CD evaluate and apply, A.p): measured best
// A.p loads its first entries before staging the halo (a multiple of kListBatchSolve). Measured
// on the 1M dam break, one run: 4 / 8 / 12 / 16 entries give 89.3 / 86.5 / 89.3 / 94.1 us per A.p.
constexpr uint32_t kListPrefetchApply = 8;
constexpr uint32_t kListBatchAp = kListBatchSolve;  // A.p's entries in flight after the prefetch
// The entries the DF/CD evaluation and the pressure apply load before their halo staging (a
// multiple of kListBatchSolve). Measured on the 1M dam break, one run: the evaluation takes
// 91 / 75 / 88 us with 2 / 4 / 8.
constexpr uint32_t kListPrefetchEval = 4;
constexpr uint32_t kListPrefetchPressure = 4;
static_assert(kListPrefetchEval % kListBatchSolve == 0 && kListPrefetchPressure % kListBatchSolve == 0,
              "the batches after the prefetch stay aligned");
static_assert(kListPrefetchApply % kListBatchAp == 0, "the batches after the prefetch stay aligned");

// Calls body(t, s) for every halo record t < H (s = its slot), with R slots per thread loaded
// ahead so the bodies' loads can be in flight together.
template <uint32_t R, typename F>
__device__ __forceinline__ void for_halo(const uint32_t* __restrict__ hs, uint32_t H, F&& body) {
    for (uint32_t t0 = threadIdx.x; t0 < H; t0 += R * blockDim.x) {
        uint32_t s[R];
#pragma unroll
        for (uint32_t q = 0; q < R; q++) {
            const uint32_t t = t0 + q * blockDim.x;
            s[q] = t < H ? hs[t] : 0u;
        }
#pragma unroll
        for (uint32_t q = 0; q < R; q++) {
            const uint32_t t = t0 + q * blockDim.x;
            if (t < H) body(t, s[q]);
        }
    }
}

// Halo count of workgroup blk: > kHaloCap (not staged) when there are no halos (hl.n null: the
// adjoint's kernels run on its taped lists).
__device__ __forceinline__ uint32_t halo_count(const Halo& hl, uint32_t blk) { return hl.n ? hl.n[blk] : kHaloCap + 1u; }

// LDS staging depth (records per workgroup, <= kHaloCap) of each halo-staged kernel: with wide
// records the depth decides how many workgroups share a CU's 160 KiB. A block with a deeper halo
// gathers from global memory instead, with the same results. On the 1M dam break:
// DF/CD evaluation, 24-byte records: 1104 -> 26.5 KB, six per CU (1264: five), 90 -> 82 us;
// CG setup, 36-byte records: 1104 -> 39.7 KB, four per CU (1264: three), 213 -> 191 us;
// A.p, 24-byte records: kHaloShallow = 1264 -> 30.3 KB, five per CU (1136, six per CU, was
// slower: its blocks over the depth cost more than a sixth workgroup gains), or kHaloCap when the
// host sees enough deeper halos (kernels.hpp).
constexpr uint32_t kStageEval = 1104;
constexpr uint32_t kStageSetup = 1104;
constexpr uint32_t kStageAp = kHaloShallow;
constexpr uint32_t kStageDensity = kHaloShallow, kStageST = kHaloShallow, kStagePApply = kHaloShallow;
static_assert(kStageEval <= kHaloCap && kStageSetup <= kHaloCap && kStageAp <= kHaloCap && kStageDensity <= kHaloCap &&
                  kStageST <= kHaloCap && kStagePApply <= kHaloCap,
              "staging depths");

// ----------------------------------------------------------------- boundary volume
// ComputeBoundaryParticleVolume_kernel (MPSSPHBaseUtil.cuh:10-41): S_b = W(0) + sum W_bk.
// S_b depends only on boundary positions, so it is computed when they change.
template <bool CUBIC>
__global__ void __launch_bounds__(kBlock) k_boundary_sum(const float4* __restrict__ bpos, uint32_t B, const uint32_t* __restrict__ nbb,
                                                       uint32_t Kbb, const uint32_t* __restrict__ cbb, float* __restrict__ S) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const float4 xi4 = bpos[i];
    const f3 xi = mk(xi4);
    const float hi = xi4.w * 4.0f;
    const float invHi = 1.0f / hi;
    float vol = Wk<CUBIC>(0.0f, invHi);
    const uint32_t c = cbb[i];
    for (uint32_t k = 0; k < c; k++) {
        const uint32_t j = nbb[size_t(i) * Kbb + k];
        vol += Wk<CUBIC>(len3(xi - mk(bpos[j])), invHi);
    }
    S[i] = vol;
}
// ComputeBoundaryParticleVolumeFinal_kernel (:78-86) on top of the AtomicAdd into the
// never-cleared volume (Q1): V <- 1/(V + S).
__global__ void __launch_bounds__(kBlock) k_boundary_volume(const float* __restrict__ S, float* __restrict__ vol, uint32_t B, int faithful) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    float v = faithful ? vol[i] : 0.0f;
    v += S[i];
    vol[i] = 1.0f / v;
}

// Several wall objects (a wall particle's object in obj[]): sums kept per object, walls of the
// own object first (ComputeBoundaryParticleVolume_kernel self + ref, MPSSPHBaseUtil.cuh:10-76).
// Per-object accumulators stay in registers: updates and reads select the element with
// unrolled compares instead of a dynamic index.
template <class T, class F>
__device__ __forceinline__ void at_obj(T (&a)[kMaxWallObjects], uint32_t o, F&& f) {
#pragma unroll
    for (uint32_t q = 0; q < kMaxWallObjects; q++)
        if (q == o) f(a[q]);
}
template <bool CUBIC>
__global__ void __launch_bounds__(kBlock) k_boundary_sum_obj(const float4* __restrict__ bpos, uint32_t B,
                                                           const uint32_t* __restrict__ nbb, uint32_t Kbb,
                                                           const uint32_t* __restrict__ cbb, const uint32_t* __restrict__ obj,
                                                           float4* __restrict__ S4) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const float4 xi4 = bpos[i];
    const f3 xi = mk(xi4);
    const float invHi = 1.0f / (xi4.w * 4.0f);
    float v[kMaxWallObjects] = {0.f, 0.f, 0.f, 0.f};
    at_obj(v, obj[i], [&](float& a) { a = Wk<CUBIC>(0.0f, invHi); });
    const uint32_t c = cbb[i];
    for (uint32_t k = 0; k < c; k++) {
        const uint32_t j = nbb[size_t(i) * Kbb + k];
        const float w = Wk<CUBIC>(len3(xi - mk(bpos[j])), invHi);
        at_obj(v, obj[j], [&](float& a) { a += w; });
    }
    S4[i] = make_float4(v[0], v[1], v[2], v[3]);
}
__global__ void __launch_bounds__(kBlock) k_boundary_volume_obj(const float4* __restrict__ S4, const uint32_t* __restrict__ obj,
                                                              uint32_t nobj, float* __restrict__ vol, uint32_t B, int faithful) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const float4 s4 = S4[i];
    const float s[kMaxWallObjects] = {s4.x, s4.y, s4.z, s4.w};
    const uint32_t own = obj[i];
    float v = faithful ? vol[i] : 0.0f;
    at_obj(s, own, [&](const float& a) { v += a; });
#pragma unroll
    for (uint32_t q = 0; q < kMaxWallObjects; q++)
        if (q < nobj && q != own) v += s[q];
    vol[i] = 1.0f / v;
}

// ----------------------------------------------------------------- density, alpha, gravity
// ComputeDensity_kernel ff/fb + Final + DensityColorTest (MPSSPHBaseUtil.cuh:88-221),
// ComputeDFSPHFactor_kernel ff/fb + Final (MPSSPHDFSPHUtil.cuh:10-137),
// ApplyGravity + UpdateVelocity (MPSAdvectUtil.cuh:8-34).
// Also stores G_ij per neighbour slot (pG_ff, pG_fb) when kDoPairs is set.
template <bool CUBIC, bool REF, bool BOBJ>
__global__ void __launch_bounds__(kBlock) k_density_alpha(int flags, Scenes sc, Nbr nb, Halo hl, const float4* __restrict__ pos,
                                                        float4* __restrict__ vel, const float4* __restrict__ bpos,
                                                        const float* __restrict__ bvol, float* __restrict__ rho,
                                                        float* __restrict__ alpha, float4* __restrict__ color,
                                                        float* __restrict__ pG_ff, float* __restrict__ pG_fb) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    const StepParams P = sc.P[sid];
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    const bool live = i < P.end;
    const bool dens = flags & launch::kDoDensity, alph = flags & launch::kDoAlpha;
    const bool pairs = flags & launch::kDoPairs;
    const bool walk = dens || alph || pairs;
    const uint32_t c = live && walk ? nb.cff[i] : 0u;
    // the block's halo positions in LDS (see k_cg_apply)
    __shared__ float4 la[kStageDensity];
    const uint32_t H = halo_count(hl, blk);
    const bool staged = walk && H <= kStageDensity;
    constexpr uint32_t U = kListBatch;
    uint32_t jn[U];
    auto fetch = [&](uint32_t k0) {
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            if (k0 + q < c) {
                const size_t slot = size_t(k0 + q) * N + i;
                jn[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
            }
        }
    };
    fetch(0);
    if (staged) {
        const uint32_t* hs = hl.idx + size_t(blk) * kHaloCap;
        for_halo<4>(hs, H, [&](uint32_t t, uint32_t s) { la[t] = pos[s]; });
        __syncthreads();
    }
    if (!live) return;
    const float4 xi4 = pos[i];
    const f3 xi = mk(xi4);
    const float hi = xi4.w * 4.0f;
    const float invHi = 1.0f / hi;
    float sf = dens ? P.V * Wk<CUBIC>(0.0f, invHi) : 0.0f;
    f3 gf{0.f, 0.f, 0.f};
    float s = 0.0f;
    if (walk) {
        auto pair = [&](uint32_t k, float4 xj4) {
            const f3 xij = xi - mk(xj4);
            const float dist = len3(xij);
            if (dens) sf += P.V * Wk<CUBIC>(dist, invHi);
            if (alph || pairs) {
                const float G = Gk<CUBIC>(dist, invHi);
                if (pairs) pG_ff[size_t(k) * N + i] = G;
                const f3 grad = (P.V * G) * xij;
                gf = gf + grad;
                s += dot3(grad, grad);
            }
        };
        for (uint32_t k = 0; k < c; k += U) {
            uint32_t j[U];
#pragma unroll
            for (uint32_t q = 0; q < U; q++) j[q] = jn[q];
            if (k + U < c) fetch(k + U);
            if (staged) {
#pragma unroll
                for (uint32_t q = 0; q < U; q++)
                    if (k + q < c) pair(k + q, la[j[q]]);
            } else {
                for (uint32_t q = 0; q < U && k + q < c; q++) pair(k + q, pos[j[q]]);
            }
        }
        float sb = 0.0f;
        f3 gb{0.f, 0.f, 0.f};
        float sbo[kMaxWallObjects] = {0.f, 0.f, 0.f, 0.f};  // per wall object (BOBJ)
        f3 gbo[kMaxWallObjects] = {};
        struct WallRec {
            float4 x;
            float vb;
            uint32_t o;
        };
        walk_walls(nb, i, N, nb.cfb[i],
                   [&](uint32_t j, size_t) { return WallRec{bpos[j], bvol[j], BOBJ ? nb.fbo[j] : 0u}; },
                   [&](uint32_t, uint32_t, size_t slot, const WallRec& w) {
            const f3 xij = xi - mk(w.x);
            const float dist = len3(xij);
            const float vb = w.vb;
            if constexpr (BOBJ) {
                if (dens) {
                    const float wk = vb * Wk<CUBIC>(dist, invHi);
                    at_obj(sbo, w.o, [&](float& a) { a += wk; });
                }
                if (alph || pairs) {
                    const float G = Gk<CUBIC>(dist, invHi);
                    if (pairs) pG_fb[slot] = G;
                    const f3 gg = (vb * G) * xij;
                    at_obj(gbo, w.o, [&](f3& a) { a = a + gg; });
                }
            } else {
                if (dens) sb += vb * Wk<CUBIC>(dist, invHi);
                if (alph || pairs) {
                    const float G = Gk<CUBIC>(dist, invHi);
                    if (pairs) pG_fb[slot] = G;
                    gb = gb + (vb * G) * xij;
                }
            }
        });
        // ((own + each ref object's sum, in scene order) + walls), as the reference's Sub
        // kernels accumulate them one after the other
        float d = 0.0f;
        d += sf;
        f3 g{0.f, 0.f, 0.f};
        g = g + gf;
        float a = 0.0f;
        a += s;
        if constexpr (REF) {  // ComputeDensity / ComputeDFSPHFactor ref (MPSSPHBaseUtil.cuh:122-156, MPSSPHDFSPHUtil.cuh:49-88)
            if (dens || alph) {
                const uint32_t cr = nb.cfr[i];
                uint32_t cur = kEmpty;
                float Vt = 0.0f, sr = 0.0f, ssr = 0.0f;
                f3 gr{0.f, 0.f, 0.f};
                for (uint32_t k = 0; k < cr; k++) {
                    const uint32_t j = nb.fr[size_t(k) * N + i];
                    const uint32_t t = sc.blk_scene[j / kBlock];
                    if (t != cur) {
                        if (cur != kEmpty) {
                            d += sr;
                            g = g + gr;
                            a += ssr;
                        }
                        cur = t;
                        Vt = sc.P[t].V;
                        sr = 0.0f;
                        ssr = 0.0f;
                        gr = f3{0.f, 0.f, 0.f};
                    }
                    const f3 xij = xi - mk(pos[j]);
                    const float dist = len3(xij);
                    if (dens) sr += Vt * Wk<CUBIC>(dist, invHi);
                    if (alph) {
                        const f3 grad = (Vt * Gk<CUBIC>(dist, invHi)) * xij;
                        gr = gr + grad;
                        ssr += dot3(grad, grad);
                    }
                }
                if (cur != kEmpty) {
                    d += sr;
                    g = g + gr;
                    a += ssr;
                }
            }
        }
        if (dens) {
            if constexpr (BOBJ) {
#pragma unroll
                for (uint32_t q = 0; q < kMaxWallObjects; q++)
                    if (q < nb.nbo) d += sbo[q];
            } else {
                d += sb;
            }
            const float r = d * P.rho0;
            rho[i] = r;
            const float ratio = r / P.rho0;
            const float blue = fminf(fmaxf(1.0f - ratio, 0.0f), 1.0f);
            const float green = fminf(fmaxf(ratio < 1.0f ? ratio : 1.0f - (ratio - 1.0f), 0.0f), 1.0f);
            const float red = fminf(fmaxf(ratio - 1.0f, 0.0f), 1.0f);
            color[i] = make_float4(red, green, blue, 1.0f);
        }
        if (alph) {
            if constexpr (BOBJ) {
#pragma unroll
                for (uint32_t q = 0; q < kMaxWallObjects; q++)
                    if (q < nb.nbo) g = g + gbo[q];
            } else {
                g = g + gb;
            }
            a += dot3(g, g);
            alpha[i] = a > kEps ? 1.0f / a : 0.0f;
        }
    }
    if (flags & launch::kDoGravity) {
        float4 v4 = vel[i];
        const float m = v4.w;
        const f3 F = f3{0.f, 0.f, 0.f} + f3{P.gx, P.gy, P.gz} * m;
        if (m > 1.0e-10f) {
            const float c = P.dt / m;
            v4.x = v4.x + c * F.x;
            v4.y = v4.y + c * F.y;
            v4.z = v4.z + c * F.z;
            vel[i] = v4;
        }
    }
}

// ----------------------------------------------------------------- pressure solves
// ComputeDensityDelta_kernel ff/fb (MPSSPHDFSPHUtil.cuh:139-257) fused with
// ComputeDFStiffness_kernel / ComputeCDStiffness_kernel (:259-340); G_ij comes from the
// alpha pass. Each workgroup publishes its sum of positive errors for k_fold_pressure.
// The records of the block's halo (x_j, v_j) are staged in LDS (see k_cg_apply); the pairs
// are summed in list order on both paths.
template <bool DF, bool REF, bool CUBIC, bool BOBJ>
__global__ void __launch_bounds__(kBlock) k_pressure_eval(Scenes sc, Nbr nb, Halo hl, const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel, const float4* __restrict__ bpos,
                                                        const float* __restrict__ bvol, const float4* __restrict__ bvel,
                                                        const float* __restrict__ rho, const float* __restrict__ alpha,
                                                        const float* __restrict__ pG_ff, const float* __restrict__ pG_fb,
                                                        float* __restrict__ kappa, float* __restrict__ delta,
                                                        const SolverDev* st, double* __restrict__ partials) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    const StepParams P = sc.P[sid];
    if (st[sid].done) return;  // this scene's loop has finished
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    const bool live = i < P.end;
    const uint32_t c = live ? nb.cff[i] : 0u;
    __shared__ float4 la[kStageEval];
    __shared__ float2 lb[kStageEval];
    const uint32_t H = hl.n[blk];
    const bool staged = H <= kStageEval;
    constexpr uint32_t U = kListBatchSolve;
    uint32_t jn[U];
    float gn[U];
    auto fetch = [&](uint32_t k0) {
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            if (k0 + q < c) {
                const size_t slot = size_t(k0 + q) * N + i;
                jn[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
                gn[q] = __builtin_nontemporal_load(pG_ff + slot);
            }
        }
    };
    // the first kPre entries are loaded before the halo staging, as in A.p
    constexpr uint32_t kPre = kListPrefetchEval;
    uint32_t jp[kPre];
    float gp[kPre];
#pragma unroll
    for (uint32_t q = 0; q < kPre; q++) {
        if (q < c) {
            const size_t slot = size_t(q) * N + i;
            jp[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
            gp[q] = __builtin_nontemporal_load(pG_ff + slot);
        }
    }
    if (staged) {
        const uint32_t* hs = hl.idx + size_t(blk) * kHaloCap;
        for_halo<4>(hs, H, [&](uint32_t t, uint32_t s) {
            const float4 x = pos[s], v = vel[s];
            la[t] = make_float4(x.x, x.y, x.z, v.x);
            lb[t] = make_float2(v.y, v.z);
        });
        __syncthreads();
    }
    double epos = 0.0;
    if (live) {
        const f3 xi = mk(pos[i]), vi = mk(vel[i]);
        float df = 0.0f;
        if (kPre < c) fetch(kPre);
#pragma unroll
        for (uint32_t q = 0; q < kPre; q++) {
            if (q < c) {
                if (staged) {
                    const float4 a = la[jp[q]];
                    const float2 b = lb[jp[q]];
                    df += P.V * gp[q] * dot3(xi - mk(a), vi - f3{a.w, b.x, b.y});
                } else {
                    df += P.V * gp[q] * dot3(xi - mk(pos[jp[q]]), vi - mk(vel[jp[q]]));
                }
            }
        }
        for (uint32_t k = kPre; k < c; k += U) {
            uint32_t j[U];
            float G[U];
#pragma unroll
            for (uint32_t q = 0; q < U; q++) {
                j[q] = jn[q];
                G[q] = gn[q];
            }
            if (k + U < c) fetch(k + U);
            if (staged) {
#pragma unroll
                for (uint32_t q = 0; q < U; q++) {
                    if (k + q < c) {
                        const float4 a = la[j[q]];
                        const float2 b = lb[j[q]];
                        df += P.V * G[q] * dot3(xi - mk(a), vi - f3{a.w, b.x, b.y});
                    }
                }
            } else {
                for (uint32_t q = 0; q < U && k + q < c; q++) df += P.V * G[q] * dot3(xi - mk(pos[j[q]]), vi - mk(vel[j[q]]));
            }
        }
        float db = 0.0f;
        float dbo[kMaxWallObjects] = {0.f, 0.f, 0.f, 0.f};
        struct WallRec {
            float4 x, v;
            float vb, G;
            uint32_t o;
        };
        walk_walls(nb, i, N, nb.cfb[i],
                   [&](uint32_t j, size_t slot) {
                       return WallRec{bpos[j], bvel[j], bvol[j], pG_fb[slot], BOBJ ? nb.fbo[j] : 0u};
                   },
                   [&](uint32_t, uint32_t, size_t, const WallRec& w) {
            const f3 xij = xi - mk(w.x);
            const f3 vij = vi - mk(w.v);
            if constexpr (BOBJ) {
                const float t = w.vb * w.G * dot3(xij, vij);
                at_obj(dbo, w.o, [&](float& a) { a += t; });
            } else {
                db += w.vb * w.G * dot3(xij, vij);
            }
        });
        float d = 0.0f;
        d += df;
        if constexpr (REF) {  // ComputeDensityDelta ref (MPSSPHDFSPHUtil.cuh:177-216): one sum per ref object
            const float invHi = 1.0f / (pos[i].w * 4.0f);
            const uint32_t cr = nb.cfr[i];
            uint32_t cur = kEmpty;
            float Vt = 0.0f, dr = 0.0f;
            for (uint32_t k = 0; k < cr; k++) {
                const uint32_t j = nb.fr[size_t(k) * N + i];
                const uint32_t t = sc.blk_scene[j / kBlock];
                if (t != cur) {
                    if (cur != kEmpty) d += dr;
                    cur = t;
                    Vt = sc.P[t].V;
                    dr = 0.0f;
                }
                const f3 xij = xi - mk(pos[j]);
                const f3 vij = vi - mk(vel[j]);
                dr += Vt * Gk<CUBIC>(len3(xij), invHi) * dot3(xij, vij);
            }
            if (cur != kEmpty) d += dr;
        }
        if constexpr (BOBJ) {
#pragma unroll
            for (uint32_t q = 0; q < kMaxWallObjects; q++)
                if (q < nb.nbo) d += dbo[q];
        } else {
            d += db;
        }
        delta[i] = d;
        float e;
        if (DF) {
            const float a = d * P.dt;
            const float b = rho[i] / P.rho0 + P.dt * d - 0.8f;
            e = a < b ? a : b;
        } else {
            e = rho[i] / P.rho0 + P.dt * d - 1.0f;
        }
        if (e > 0.0f) {
            epos = double(e);
            e *= alpha[i] / (P.dt * P.dt);
        } else {
            e = 0.0f;
        }
        kappa[i] = e;
    }
    block_partial(epos, partials, blk);
}

// Mean positive error and the loop decision of the DF/CD drivers
// (MPSSPHDFSPHUtil.cu:331-333,417-419): break when E < tol and l > lmin. Workgroup s folds
// scene s's partials; a scene that finishes leaves the count of active scenes.
// The iteration number is the record's l (1 after the reset, it + 1 after iteration it), so the
// launches of a chunk do not depend on where the chunk starts.
__global__ void __launch_bounds__(kFoldThreads) k_fold_pressure(Scenes sc, const double* __restrict__ partials,
                                                        SolverDev* st, uint32_t* active, double tol, uint32_t lmin) {
    const uint32_t sid = blockIdx.x;
    SolverDev* sd = st + sid;
    if (sd->done) return;
    const uint32_t it = sd->l;
    const StepParams P = sc.P[sid];
    const double s = fold(partials + P.blk0, P.nblk);
    if (threadIdx.x == 0) {
        const double E = s / double(P.count);
        if (it - 1 < MPS_MAX_ITER_LOG) sd->log[it - 1] = E;
        if (E < tol && it > lmin) {
            sd->done = 1u;
            sd->l = it;
            atomicSub(active, 1u);
        } else {
            sd->l = it + 1;
        }
    }
}

// Coupled fluid objects run one pressure loop: a single workgroup folds every scene's errors
// and the loop ends at the first iteration where every scene has E < tol and l > lmin (all
// scenes then finish together, with the same l).
__global__ void __launch_bounds__(kFoldThreads) k_fold_pressure_joint(Scenes sc, const double* __restrict__ partials,
                                                              SolverDev* st, uint32_t* active, double tol, uint32_t lmin) {
    if (st[0].done) return;
    const uint32_t it = st[0].l;
    __shared__ uint32_t all_ok;
    if (threadIdx.x == 0) all_ok = 1u;
    __syncthreads();
    for (uint32_t sid = 0; sid < sc.S; sid++) {
        const StepParams P = sc.P[sid];
        const double s = fold(partials + P.blk0, P.nblk);
        if (threadIdx.x == 0) {
            const double E = P.count ? s / double(P.count) : 0.0;
            if (it - 1 < MPS_MAX_ITER_LOG) st[sid].log[it - 1] = E;
            if (!(E < tol && it > lmin)) all_ok = 0u;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const bool ok = all_ok != 0u;
        for (uint32_t sid = 0; sid < sc.S; sid++) {
            if (ok) st[sid].done = 1u;
            st[sid].l = ok ? it : it + 1;
        }
        if (ok) *active = 0u;
    }
}

// ApplyDFSPH_kernel ff/fb + ApplyDFSPHFinal_kernel (MPSSPHDFSPHUtil.cuh:342-468) + UpdateVelocity.
// The halo's (x_j, kappa_j) are staged in LDS (see k_cg_apply).
template <bool REF, bool CUBIC, bool BOBJ>
__global__ void __launch_bounds__(kBlock) k_pressure_apply(Scenes sc, Nbr nb, Halo hl, const float4* __restrict__ pos,
                                                         float4* __restrict__ vel, const float4* __restrict__ bpos,
                                                         const float* __restrict__ bvol, const float* __restrict__ pG_ff,
                                                         const float* __restrict__ pG_fb, const float* __restrict__ kappa,
                                                         const SolverDev* st) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    const StepParams P = sc.P[sid];
    if (st[sid].done) return;
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    const bool live = i < P.end;
    const uint32_t c = live ? nb.cff[i] : 0u;
    __shared__ float4 la[kStagePApply];
    const uint32_t H = hl.n[blk];
    const bool staged = H <= kStagePApply;
    constexpr uint32_t U = kListBatchSolve;
    uint32_t jn[U];
    float gn[U];
    auto fetch = [&](uint32_t k0) {
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            if (k0 + q < c) {
                const size_t slot = size_t(k0 + q) * N + i;
                jn[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
                gn[q] = __builtin_nontemporal_load(pG_ff + slot);
            }
        }
    };
    // the first kPre entries are loaded before the halo staging, as in A.p
    constexpr uint32_t kPre = kListPrefetchPressure;
    uint32_t jp[kPre];
    float gp[kPre];
#pragma unroll
    for (uint32_t q = 0; q < kPre; q++) {
        if (q < c) {
            const size_t slot = size_t(q) * N + i;
            jp[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
            gp[q] = __builtin_nontemporal_load(pG_ff + slot);
        }
    }
    if (staged) {
        const uint32_t* hs = hl.idx + size_t(blk) * kHaloCap;
        for_halo<4>(hs, H, [&](uint32_t t, uint32_t s) {
            const float4 x = pos[s];
            la[t] = make_float4(x.x, x.y, x.z, kappa[s]);
        });
        __syncthreads();
    }
    if (!live) return;
    const f3 xi = mk(pos[i]);
    const float pi = kappa[i];
    f3 ff{0.f, 0.f, 0.f};
    if (kPre < c) fetch(kPre);
#pragma unroll
    for (uint32_t q = 0; q < kPre; q++) {
        if (q < c) {
            if (staged) {
                const float4 a = la[jp[q]];
                ff = ff - 0.5f * P.V * (pi + a.w) * gp[q] * (xi - mk(a));
            } else {
                ff = ff - 0.5f * P.V * (pi + kappa[jp[q]]) * gp[q] * (xi - mk(pos[jp[q]]));
            }
        }
    }
    for (uint32_t k = kPre; k < c; k += U) {
        uint32_t j[U];
        float G[U];
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            j[q] = jn[q];
            G[q] = gn[q];
        }
        if (k + U < c) fetch(k + U);
        if (staged) {
#pragma unroll
            for (uint32_t q = 0; q < U; q++) {
                if (k + q < c) {
                    const float4 a = la[j[q]];
                    ff = ff - 0.5f * P.V * (pi + a.w) * G[q] * (xi - mk(a));
                }
            }
        } else {
            for (uint32_t q = 0; q < U && k + q < c; q++) ff = ff - 0.5f * P.V * (pi + kappa[j[q]]) * G[q] * (xi - mk(pos[j[q]]));
        }
    }
    f3 fb{0.f, 0.f, 0.f};
    f3 fbo[kMaxWallObjects] = {};
    struct WallRec {
        float4 x;
        float vb, G;
        uint32_t o;
    };
    walk_walls(nb, i, N, nb.cfb[i],
               [&](uint32_t j, size_t slot) { return WallRec{bpos[j], bvol[j], pG_fb[slot], BOBJ ? nb.fbo[j] : 0u}; },
               [&](uint32_t, uint32_t, size_t, const WallRec& w) {
        const f3 xij = xi - mk(w.x);
        const f3 fij = 0.5f * w.vb * pi * w.G * xij;
        if constexpr (BOBJ) at_obj(fbo, w.o, [&](f3& a) { a = a - fij; });
        else fb = fb - fij;
    });
    f3 F{0.f, 0.f, 0.f};
    F = F + ff;
    if constexpr (REF) {  // ApplyDFSPH ref (MPSSPHDFSPHUtil.cuh:379-418): one sum per ref object
        const float invHi = 1.0f / (pos[i].w * 4.0f);
        const uint32_t cr = nb.cfr[i];
        uint32_t cur = kEmpty;
        float Vt = 0.0f;
        f3 fr{0.f, 0.f, 0.f};
        for (uint32_t k = 0; k < cr; k++) {
            const uint32_t j = nb.fr[size_t(k) * N + i];
            const uint32_t t = sc.blk_scene[j / kBlock];
            if (t != cur) {
                if (cur != kEmpty) F = F + fr;
                cur = t;
                Vt = sc.P[t].V;
                fr = f3{0.f, 0.f, 0.f};
            }
            const f3 xij = xi - mk(pos[j]);
            fr = fr - 0.5f * Vt * (pi + kappa[j]) * Gk<CUBIC>(len3(xij), invHi) * xij;
        }
        if (cur != kEmpty) F = F + fr;
    }
    if constexpr (BOBJ) {
#pragma unroll
        for (uint32_t q = 0; q < kMaxWallObjects; q++)
            if (q < nb.nbo) F = F + fbo[q];
    } else {
        F = F + fb;
    }
    float4 v4 = vel[i];
    const float m = v4.w;
    F = m * F;
    if (m > 1.0e-10f) {
        const float cdt = P.dt / m;
        v4.x = v4.x + cdt * F.x;
        v4.y = v4.y + cdt * F.y;
        v4.z = v4.z + cdt * F.z;
        vel[i] = v4;
    }
}

// ----------------------------------------------------------------- viscosity + surface tension CG
// ComputeSurfaceTensionFactor_kernel (MPSSPHNonPressureUtil.cuh:425-465) with the stored
// G_ij; also x0 = v (the copy into pPredictVel, MPSSPHNonPressureUtil.cu:144-147).
__global__ void __launch_bounds__(kBlock) k_st_factor(Scenes sc, Nbr nb, Halo hl, const float4* __restrict__ pos,
                                                    const float4* __restrict__ vel, const float* __restrict__ pG_ff,
                                                    float* __restrict__ fst, float4* __restrict__ cgx) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    const StepParams P = sc.P[sid];
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    const bool live = i < P.end;
    const uint32_t c = live ? nb.cff[i] : 0u;
    // the block's halo (x_j, v_j, m_j) in LDS (see k_cg_apply)
    __shared__ float4 la[kStageST], lb[kStageST];
    const uint32_t H = halo_count(hl, blk);
    const bool staged = H <= kStageST;
    constexpr uint32_t U = kListBatch;
    uint32_t jn[U];
    float gn[U];
    auto fetch = [&](uint32_t k0) {
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            if (k0 + q < c) {
                const size_t slot = size_t(k0 + q) * N + i;
                jn[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
                gn[q] = __builtin_nontemporal_load(pG_ff + slot);
            }
        }
    };
    fetch(0);
    if (staged) {
        const uint32_t* hs = hl.idx + size_t(blk) * kHaloCap;
        for_halo<4>(hs, H, [&](uint32_t t, uint32_t s) {
            const float4 x = pos[s], v = vel[s];
            la[t] = make_float4(x.x, x.y, x.z, v.x);
            lb[t] = make_float4(v.y, v.z, v.w, 0.0f);
        });
        __syncthreads();
    }
    if (!live) return;
    const float4 vi4 = vel[i];
    const f3 xi = mk(pos[i]), vi = mk(vi4);
    float d = 0.0f;
    for (uint32_t k = 0; k < c; k += U) {
        uint32_t j[U];
        float G[U];
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            j[q] = jn[q];
            G[q] = gn[q];
        }
        if (k + U < c) fetch(k + U);
        if (staged) {
#pragma unroll
            for (uint32_t q = 0; q < U; q++) {
                if (k + q < c) {
                    const float4 e = la[j[q]], g = lb[j[q]];
                    d += g.z * dot3(vi - f3{e.w, g.x, g.y}, G[q] * (xi - mk(e)));
                }
            }
        } else {
            for (uint32_t q = 0; q < U && k + q < c; q++) {
                const float4 vj = vel[j[q]];
                d += vj.w * dot3(vi - mk(vj), G[q] * (xi - mk(pos[j[q]])));
            }
        }
    }
    fst[i] = d;
    cgx[i] = make_float4(vi4.x, vi4.y, vi4.z, 0.0f);
}

// g_ij of the surface-tension operator (MPSSPHNonPressureUtil.cuh:503-512,563-572)
__device__ __forceinline__ float st_g(const StepParams& P, f3 vij, f3 xij, float dist, float invHi, float mi, float mj,
                                      float di, float dj, float fi, float fj, float& mij) {
    mij = 0.5f * (mi + mj);
    const float invDij = 2.0f / (di + dj);
    const float stW = STWk(dist, invHi);
    const f3 stG = STGk(dist, invHi) * xij;
    return (stW + P.dt * (dot3(vij, stG) - (fi + fj) * invDij * 0.5f * stW)) * invDij;
}

// Once per solve: right-hand side (ComputeSurfaceTensionCGb_kernel, :467-521) and the pair
// coefficients of the implicit operator (ST :523-585, viscosity :587-635, final :637-648):
//   a_ij = dt (cs_i m_ij g_ij + cx (m_j/rho_j) W_ij),   b_ij = -dt cv_i c_ij,
//   c_ij = (0.5(m_i+m_j) (1/rho_j)) Phi(d_ij)   (the reference's (0.5(m_i+m_j)/rho_j) /
//          (d^2 + 0.01 h_i^2) G_ij / (d + FLT_EPSILON) regrouped, see Phik),
//   cs_i = dt sigma/m_i, cx = XSPH/dt, cv_i = 10 rho0 mu/rho_i.
// Writes a_ij (ca) and/or (a_ij, b_ij) (cab), and 1/rho_i per slot (irho) for the A.p that
// recomputes b_ij.
template <bool CUBIC>
__global__ void __launch_bounds__(kBlock) k_cg_setup(Scenes sc, Nbr nb, Halo hl, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const float* __restrict__ rho,
                                                   const float* __restrict__ fst, float* __restrict__ ca, float2* __restrict__ cab, float4* __restrict__ cgb,
                                                   float* __restrict__ irho) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    const StepParams P = sc.P[sid];
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    const bool live = i < P.end;
    const uint32_t c = live ? nb.cff[i] : 0u;
    // the block's halo (x_j, v_j, m_j, rho_j, f_j) in LDS (see k_cg_apply)
    __shared__ float4 la[kStageSetup], lb[kStageSetup];
    __shared__ float lc[kStageSetup];
    const uint32_t H = halo_count(hl, blk);
    const bool staged = H <= kStageSetup;
    constexpr uint32_t U = kListBatch;
    uint32_t jn[U];
    auto fetch = [&](uint32_t k0) {
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            if (k0 + q < c) {
                const size_t slot = size_t(k0 + q) * N + i;
                jn[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
            }
        }
    };
    fetch(0);
    if (staged) {
        const uint32_t* hs = hl.idx + size_t(blk) * kHaloCap;
        for_halo<4>(hs, H, [&](uint32_t t, uint32_t s) {
            const float4 x = pos[s], v = vel[s];
            la[t] = make_float4(x.x, x.y, x.z, v.x);
            lb[t] = make_float4(v.y, v.z, v.w, rho[s]);
            lc[t] = fst[s];
        });
        __syncthreads();
    }
    if (!live) return;
    const float4 xi4 = pos[i];
    const float4 vi4 = vel[i];
    const f3 xi = mk(xi4), vi = mk(vi4);
    const float mi = vi4.w, di = rho[i], fi = fst[i];
    const float hi = xi4.w * 4.0f;
    const float invHi = 1.0f / hi;
    const float onePercentHSq = 0.01f * hi * hi;
    const float cs = P.dt * P.sigma / mi;
    const float cx = P.xsph / P.dt;
    const float cv = 10.0f * P.rho0 * P.mu / di;
    if (irho) irho[i] = 1.0f / di;
    f3 bi{0.f, 0.f, 0.f};
    auto pair = [&](size_t slot, f3 xj, f3 vj, float mj, float dj, float fj) {
        const f3 xij = xi - xj;
        const float dist = len3(xij);
        float mij;
        const float gij = st_g(P, vi - vj, xij, dist, invHi, mi, mj, di, dj, fi, fj, mij);
        const float sij = mij * gij;
        const float wij = mj / dj * Wk<CUBIC>(dist, invHi);
        const float aij = P.dt * (cs * sij + cx * wij);
        if (ca) ca[slot] = aij;
        if (cab) {
            const float irhoj = 1.0f / dj;
            const float cij = 0.5f * (mi + mj) * irhoj * Phik<CUBIC>(dist, invHi, onePercentHSq);
            cab[slot] = make_float2(aij, -P.dt * cv * cij);
        }
        bi = bi - sij * xij;
    };
    for (uint32_t k = 0; k < c; k += U) {
        uint32_t j[U];
#pragma unroll
        for (uint32_t q = 0; q < U; q++) j[q] = jn[q];
        if (k + U < c) fetch(k + U);
        if (staged) {
#pragma unroll
            for (uint32_t q = 0; q < U; q++) {
                if (k + q < c) {
                    const float4 e = la[j[q]], g = lb[j[q]];
                    pair(size_t(k + q) * N + i, mk(e), f3{e.w, g.x, g.y}, g.z, g.w, lc[j[q]]);
                }
            }
        } else {
            for (uint32_t q = 0; q < U && k + q < c; q++) {
                const uint32_t jj = j[q];
                const float4 vj = vel[jj];
                pair(size_t(k + q) * N + i, mk(pos[jj]), mk(vj), vj.w, rho[jj], fst[jj]);
            }
        }
    }
    const f3 b = (P.dt * P.sigma / mi) * bi + vi;
    cgb[i] = make_float4(b.x, b.y, b.z, 0.0f);
}

// CG vectors that neighbours read are kept in 32-byte records with the position, so one
// gather per pair brings both x_j and p_j (one 32-byte sector): rec[2j] = (x, y, z, p.x),
// rec[2j+1] = (p.y, p.z, 0, 0).
__device__ __forceinline__ f3 rec_p(float4 a, float4 b) { return {a.w, b.x, b.y}; }
__device__ __forceinline__ f3 rec_p(float4 a, float2 b) { return {a.w, b.x, b.y}; }
__global__ void __launch_bounds__(kBlock) k_cg_pack(uint32_t N, const float4* __restrict__ pos, const float4* __restrict__ x0,
                                                  float4* __restrict__ rec) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float4 x = pos[i], v = x0[i];
    rec[2 * size_t(i)] = make_float4(x.x, x.y, x.z, v.x);
    rec[2 * size_t(i) + 1] = make_float4(v.y, v.z, 0.0f, 0.0f);
}

// (A p)_i = (1+f) p_i + sum_j [ a_ij p_ij + (b_ij (p_ij . x_ij)) x_ij ].
// a_ij is written once per solve by k_cg_setup and streamed here with non-temporal loads. b_ij
// is either streamed with it (RECOMP false: 8 bytes per pair) or recomputed here from the
// geometry (RECOMP true: 4 bytes per pair streamed, one square root and one division per pair:
// b_ij = (-dt cv_i) ((0.5 (m_i + m_j) (1/rho_j)) Phi(d_ij)), the expression k_cg_setup
// evaluates). The block's halo records (x_j, p_j: 24 bytes; with m_j and 1/rho_j when b_ij is
// recomputed: 32 bytes) are staged in LDS first, with coalesced loads of the sorted halo list,
// and each pair reads its neighbour there through the 16-bit list; a block whose halo is over
// the limit gathers from global memory. Both paths visit the same pairs in the same order, so
// the result depends on neither the path nor RECOMP.
// The search direction of UpdateCGDirection_kernel (MPSCGKernel.cuh:123-133),
// p = r + beta p_prev, is formed here for every record read (own and halo) instead of in a
// pass of its own; the block writes its own p for the update kernel (pout, while pprev is
// read: the host alternates the two buffers). beta = 0 in the first iteration, where
// pprev = r, so p = r.
// The CG vectors are packed 12-byte triples (f3).
// MODE 0: A x0 with f = 0 (x0 in x0) fused with InitCGResidual_kernel (MPSCGKernel.cuh:5-19):
//         r = b - A x0, pout = r, x = x0, partial sums of r.r.
// MODE 1: A p with f (r in rr), partial sums of p.Ap (MPSCGKernel.cuh:21-67).
// One workgroup's share of A.p: block `blk` of scene `sid`, search-direction factor bta.
// PCOH: the partial sum is stored device-coherently (a workgroup of the same launch folds it).
template <int MODE, bool RECOMP, bool CUBIC, bool COH = false, bool PCOH = COH, uint32_t DEPTH = kStageAp>
__device__ __forceinline__ void cg_apply_block(uint32_t blk, uint32_t sid, float bta, const Scenes& sc, const Nbr& nb,
                                               const Halo& hl, const float2* __restrict__ cab, const float* __restrict__ ca,
                                               const float4* __restrict__ pos, const float4* __restrict__ vel,
                                               const float* __restrict__ rho, const float* __restrict__ irho,
                                               const float4* __restrict__ x0, const f3* __restrict__ rr,
                                               const f3* __restrict__ pprev, f3* __restrict__ pout, f3* __restrict__ Ap,
                                               const float4* __restrict__ b, f3* __restrict__ r, f3* __restrict__ x,
                                               double* __restrict__ partials) {
    const StepParams P = sc.P[sid];
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    const bool live = i < P.end;
    const uint32_t c = live ? nb.cff[i] : 0u;
    // the vector the operator is applied to, at slot j
    auto vec = [&](uint32_t j) -> f3 {
        if (MODE == 0) return mk(x0[j]);
        const f3 a = ld3c<COH>(rr + j), q = ld3c<COH>(pprev + j);
        return {a.x + bta * q.x, a.y + bta * q.y, a.z + bta * q.z};
    };
    using LB = typename std::conditional<RECOMP, float4, float2>::type;
    __shared__ float4 la[DEPTH];
    __shared__ LB lb[DEPTH];
    const uint32_t H = hl.n[blk];
    const bool staged = H <= DEPTH;  // same for the whole block
    using Coef = typename std::conditional<RECOMP, float, float2>::type;
    auto coef = [&](size_t slot) -> Coef {
        if constexpr (RECOMP) {
            return __builtin_nontemporal_load(ca + slot);
        } else {
            const double d = __builtin_nontemporal_load(reinterpret_cast<const double*>(cab + slot));
            return __builtin_bit_cast(float2, d);
        }
    };
    constexpr uint32_t U = kListBatchAp;
    uint32_t jn[U];
    Coef abn[U];
    auto fetch = [&](uint32_t k0) {
#pragma unroll
        for (uint32_t q = 0; q < U; q++) {
            if (k0 + q < c) {
                const size_t slot = size_t(k0 + q) * N + i;
                jn[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
                abn[q] = coef(slot);
            }
        }
    };
    // The first kListPrefetchApply entries are loaded before the halo staging: their stream
    // overlaps the staging's gathers instead of following them.
    constexpr uint32_t kPre = kListPrefetchApply;
    uint32_t jp[kPre];
    Coef abp[kPre];
#pragma unroll
    for (uint32_t q = 0; q < kPre; q++) {
        if (q < c) {
            const size_t slot = size_t(q) * N + i;
            jp[q] = staged ? uint32_t(hl.nb16[slot]) : __builtin_nontemporal_load(nb.ff + slot);
            abp[q] = coef(slot);
        }
    }
    if (staged) {
        const uint32_t* hs = hl.idx + size_t(blk) * kHaloCap;
        for_halo<4>(hs, H, [&](uint32_t t, uint32_t s) {
            const float4 xs = pos[s];
            const f3 v = vec(s);
            la[t] = make_float4(xs.x, xs.y, xs.z, v.x);
            if constexpr (RECOMP) lb[t] = make_float4(v.y, v.z, vel[s].w, irho[s]);
            else lb[t] = make_float2(v.y, v.z);
        });
        __syncthreads();
    }
    double part = 0.0;
    if (live) {
        const float4 xi4 = pos[i];
        const f3 xi = mk(xi4);
        const f3 pi = vec(i);
        // b_ij's per-particle factors, as k_cg_setup forms them
        float mi = 0.0f, ncv = 0.0f, invHi = 0.0f, h2 = 0.0f;
        if constexpr (RECOMP) {
            mi = vel[i].w;
            const float cv = 10.0f * P.rho0 * P.mu / rho[i];
            ncv = -P.dt * cv;
            const float hi = xi4.w * 4.0f;
            invHi = 1.0f / hi;
            h2 = 0.01f * hi * hi;
        }
        f3 acc{0.f, 0.f, 0.f};
        // one pair: record (xj, pj) and, for RECOMP, (mj, 1/rho_j)
        auto pair = [&](Coef ab, f3 xj, f3 pj, float mj, float irhoj) {
            const f3 xij = xi - xj;
            const f3 pij = pi - pj;
            if constexpr (RECOMP) {
                const float dist = len3(xij);
                const float cij = 0.5f * (mi + mj) * irhoj * Phik<CUBIC>(dist, invHi, h2);
                const float bij = ncv * cij;
                acc = acc + (ab * pij + (bij * dot3(pij, xij)) * xij);
            } else {
                (void)mj;
                (void)irhoj;
                acc = acc + (ab.x * pij + (ab.y * dot3(pij, xij)) * xij);
            }
        };
        auto pair_lds = [&](Coef ab, uint32_t t) {
            const float4 a = la[t];
            const LB e = lb[t];
            if constexpr (RECOMP) pair(ab, mk(a), f3{a.w, e.x, e.y}, e.z, e.w);
            else pair(ab, mk(a), f3{a.w, e.x, e.y}, 0.0f, 0.0f);
        };
        auto pair_glb = [&](Coef ab, uint32_t j) {
            if constexpr (RECOMP) pair(ab, mk(pos[j]), vec(j), vel[j].w, irho[j]);
            else pair(ab, mk(pos[j]), vec(j), 0.0f, 0.0f);
        };
        if (kPre < c) fetch(kPre);
        // the prefetched entries, then batches, all in list order
#pragma unroll
        for (uint32_t q = 0; q < kPre; q++) {
            if (q < c) {
                if (staged) pair_lds(abp[q], jp[q]);
                else pair_glb(abp[q], jp[q]);
            }
        }
        for (uint32_t k = kPre; k < c; k += U) {
            uint32_t j[U];
            Coef ab[U];
#pragma unroll
            for (uint32_t q = 0; q < U; q++) {
                j[q] = jn[q];
                ab[q] = abn[q];
            }
            if (k + U < c) fetch(k + U);
            if (staged) {
#pragma unroll
                for (uint32_t q = 0; q < U; q++)
                    if (k + q < c) pair_lds(ab[q], j[q]);
            } else {  // halo over the limit: gather from global memory
                for (uint32_t q = 0; q < U && k + q < c; q++) pair_glb(ab[q], j[q]);
            }
        }
        const float f = MODE == 0 ? 0.0f : P.cg_factor;
        const f3 o = pi * (1.0f + f) + acc;
        if (MODE == 0) {
            const float4 bi = b[i];
            const f3 ri{bi.x - o.x, bi.y - o.y, bi.z - o.z};
            r[i] = ri;
            pout[i] = ri;
            x[i] = pi;
            part = double(ri.x * ri.x) + double(ri.y * ri.y) + double(ri.z * ri.z);
        } else {
            st3c<COH>(pout + i, pi);
            Ap[i] = o;
            part = double(pi.x * o.x) + double(pi.y * o.y) + double(pi.z * o.z);
        }
    }
    block_partial<PCOH>(part, partials, blk);
}

template <int MODE, bool RECOMP, bool CUBIC, bool DEEP = false>
__global__ void __launch_bounds__(kBlock) k_cg_apply(Scenes sc, Nbr nb, Halo hl, const float2* __restrict__ cab,
                                                   const float* __restrict__ ca, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const float* __restrict__ rho,
                                                   const float* __restrict__ irho, const float4* __restrict__ x0,
                                                   const f3* __restrict__ rr, const f3* __restrict__ pprev,
                                                   f3* __restrict__ pout, f3* __restrict__ Ap,
                                                   const float4* __restrict__ b, f3* __restrict__ r, f3* __restrict__ x,
                                                   const SolverDev* st, double* __restrict__ partials) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    if (MODE == 1 && st[sid].done) return;
    const float bta = MODE == 1 ? st[sid].beta : 0.0f;
    cg_apply_block<MODE, RECOMP, CUBIC, false, false, DEEP ? kHaloCap : kStageAp>(
        blk, sid, bta, sc, nb, hl, cab, ca, pos, vel, rho, irho, x0, rr, pprev, pout, Ap, b, r, x, partials);
}

// Scalars of CGSolver (MPSSolver.cuh:31-65). MODE 0: r.r of the initial residual and the
// first convergence test; 1: alpha = r.r / p.Ap; 2: r.r of the new residual, beta, test.
// Workgroup s folds scene s's partials.
// MODE 2 takes the iteration number from the record's l (it, when iteration it is folded).
template <int MODE>
__global__ void __launch_bounds__(kFoldThreads) k_fold_cg(Scenes sc, const double* __restrict__ partials, SolverDev* stv,
                                                  uint32_t* active, double tol) {
    const uint32_t sid = blockIdx.x;
    SolverDev* st = stv + sid;
    if (MODE != 0 && st->done) return;
    const uint32_t it = st->l;
    const StepParams P = sc.P[sid];
    const double s = fold(partials + P.blk0, P.nblk);
    if (threadIdx.x != 0) return;
    if (MODE == 0) {
        st->rr = s;
        st->log[0] = s;
        st->l = 1u;
        st->thr = tol < 0.0 ? -tol * s : tol;  // a negative tolerance is relative to r0.r0
        if (s < st->thr || (tol < 0.0 && s == 0.0)) {
            st->done = 1u;
            atomicSub(active, 1u);
        }
    } else if (MODE == 1) {
        st->pAp = s;
        st->alpha = float(st->rr / s);
    } else {
        st->beta = float(s / st->rr);
        st->rr = s;
        if (it < MPS_MAX_ITER_LOG) st->log[it] = s;
        st->l = it + 1;
        if (s < st->thr || (tol < 0.0 && s == 0.0)) {
            st->done = 1u;
            atomicSub(active, 1u);
        }
    }
}

// UpdateCGResidual_kernel + the r.r half of ComputeCGBetaParam_kernel (MPSCGKernel.cuh:69-121).
// p in the 32-byte records (the adjoint's transposed solve).
__global__ void __launch_bounds__(kBlock) k_cg_update(Scenes sc, float4* __restrict__ x, float4* __restrict__ r,
                                                    const float4* __restrict__ prec, const float4* __restrict__ Ap,
                                                    const SolverDev* st, double* __restrict__ partials) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    if (st[sid].done) return;
    const uint32_t end = sc.P[sid].end;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const float a = st[sid].alpha;
    double part = 0.0;
    if (i < end) {
        const f3 pi = rec_p(prec[2 * size_t(i)], prec[2 * size_t(i) + 1]);
        const float4 api = Ap[i];
        float4 xi = x[i], ri = r[i];
        xi.x = xi.x + a * pi.x;
        xi.y = xi.y + a * pi.y;
        xi.z = xi.z + a * pi.z;
        ri.x = ri.x - a * api.x;
        ri.y = ri.y - a * api.y;
        ri.z = ri.z - a * api.z;
        x[i] = xi;
        r[i] = ri;
        part = double(ri.x * ri.x) + double(ri.y * ri.y) + double(ri.z * ri.z);
    }
    block_partial(part, partials);
}

// The update on packed triples (the forward solve): same expressions as k_cg_update.
template <bool COH = false, bool PCOH = COH>
__device__ __forceinline__ void cg_update3_block(uint32_t blk, uint32_t end, float a, f3* __restrict__ x,
                                                 f3* __restrict__ r, const f3* __restrict__ p, const f3* __restrict__ Ap,
                                                 double* __restrict__ partials) {
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < end) {
        const f3 pi = ld3c<COH>(p + i), api = Ap[i];
        f3 xi = x[i], ri = ld3c<COH>(r + i);
        xi.x = xi.x + a * pi.x;
        xi.y = xi.y + a * pi.y;
        xi.z = xi.z + a * pi.z;
        ri.x = ri.x - a * api.x;
        ri.y = ri.y - a * api.y;
        ri.z = ri.z - a * api.z;
        x[i] = xi;
        st3c<COH>(r + i, ri);
        part = double(ri.x * ri.x) + double(ri.y * ri.y) + double(ri.z * ri.z);
    }
    block_partial<PCOH>(part, partials, blk);
}
__global__ void __launch_bounds__(kBlock) k_cg_update3(Scenes sc, f3* __restrict__ x, f3* __restrict__ r,
                                                     const f3* __restrict__ p, const f3* __restrict__ Ap,
                                                     const SolverDev* st, double* __restrict__ partials) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    if (st[sid].done) return;
    cg_update3_block(blockIdx.x, sc.P[sid].end, st[sid].alpha, x, r, p, Ap, partials);
}

// ----------------------------------------------------------------- persistent CG
// The fold of the kFoldThreads-thread fold kernels, computed by one kBlock-thread workgroup in
// exactly their order: thread t plays their threads t + kBlock v (v < kFoldThreads / kBlock);
// each sums its strided partials left to right, each of their waves is reduced by the same
// butterfly (the played threads of one of their waves all sit in one wave here, in the same
// lanes), and thread 0 adds the wave totals in wave order. Valid in every thread.
template <bool COH>
__device__ __forceinline__ double fold_as_fold_kernel(const double* __restrict__ partials, uint32_t n) {
    constexpr uint32_t V = kFoldThreads / kBlock;
    constexpr uint32_t kWaves = kFoldThreads / 64;
    constexpr uint32_t R = 4;  // strides of kFoldThreads loaded per batch: R * V loads in flight
    __shared__ double wsum[kWaves];
    __shared__ double total;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    auto ld = [&](uint32_t k) -> double {
        if constexpr (COH) return __hip_atomic_load(partials + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else return partials[k];
    };
    double s[V];
#pragma unroll
    for (uint32_t v = 0; v < V; v++) s[v] = 0.0;
    for (uint32_t m0 = 0; m0 * kFoldThreads < n; m0 += R) {
        double t[R][V];
#pragma unroll
        for (uint32_t r = 0; r < R; r++)
#pragma unroll
            for (uint32_t v = 0; v < V; v++) {
                const uint32_t k = threadIdx.x + v * kBlock + (m0 + r) * kFoldThreads;
                t[r][v] = k < n ? ld(k) : 0.0;
            }
#pragma unroll
        for (uint32_t r = 0; r < R; r++)
#pragma unroll
            for (uint32_t v = 0; v < V; v++)
                if (threadIdx.x + v * kBlock + (m0 + r) * kFoldThreads < n) s[v] += t[r][v];
    }
#pragma unroll
    for (uint32_t v = 0; v < V; v++) {
        double x = s[v];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) wsum[wave + v * (kBlock / 64)] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t w = 0; w < kWaves; w++) t += wsum[w];
        total = t;
    }
    __syncthreads();
    return total;
}

// Grid-wide barrier of a co-resident grid (cooperative launch). The wait is bounded: after
// `spin` polls, or when another workgroup gave up, it sets gs->gave_up and returns false, and
// the caller leaves the kernel; every workgroup's wait then ends within `spin` polls.
// COH false: agent-scope fences publish the workgroup's writes before it arrives and make the
// others' visible after the barrier (each workgroup writes back and invalidates its L2).
// COH true: the data that crosses workgroups is written and read device-coherently (ld3c,
// st3c, coherent partials), so each wave only waits until its own stores are performed before
// the workgroup arrives, and the barrier orders nothing else.
template <bool COH>
__device__ __forceinline__ bool grid_barrier(GridSync* gs, uint32_t nwg, uint32_t spin) {
    __shared__ uint32_t ok_s;
    if constexpr (COH) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1u;
        const uint32_t g = __hip_atomic_load(&gs->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if constexpr (!COH) __threadfence();
        // arrivals are counted per group of kGridGroup workgroups, and the last of each group
        // counts the group: 16 + nwg/16 serialized atomics instead of nwg
        const uint32_t grp = blockIdx.x / kGridGroup;
        const uint32_t gsize = min(kGridGroup, nwg - grp * kGridGroup);
        const uint32_t ngrp = (nwg + kGridGroup - 1u) / kGridGroup;
        bool last = false;
        if (__hip_atomic_fetch_add(&gs->group[grp], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == gsize) {
            __hip_atomic_store(&gs->group[grp], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if constexpr (COH) __builtin_amdgcn_s_waitcnt(0);
            if (__hip_atomic_fetch_add(&gs->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == ngrp) {
                last = true;
                __hip_atomic_store(&gs->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if constexpr (COH) {
                    __builtin_amdgcn_s_waitcnt(0);
                    __hip_atomic_fetch_add(&gs->gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    __hip_atomic_fetch_add(&gs->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        if (!last) {
            uint32_t polls = 0;
            while (__hip_atomic_load(&gs->gen, COH ? __ATOMIC_RELAXED : __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                if (__hip_atomic_load(&gs->gave_up, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || ++polls > spin) {
                    ok = 0u;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        if (!ok) __hip_atomic_store(&gs->gave_up, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if constexpr (!COH) __threadfence();
        ok_s = ok && __hip_atomic_load(&gs->gave_up, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
    }
    __syncthreads();
    return ok_s != 0u;
}

// The CG iterations of CGSolver (MPSSolver.cuh:31-65) in one launch of a co-resident grid (one
// workgroup per 256-slot block of every scene), for iterations 1..cap after the initial
// residual (k_cg_apply<0> + k_fold_cg<0> ran before). Per iteration: A.p with the direction
// update (cg_apply_block) -> barrier -> every workgroup folds p.Ap itself (the fold kernels'
// order, so alpha has the same bits) -> x, r update (cg_update3_block) -> barrier -> r.r fold,
// beta and the test. The scene's first workgroup keeps the solver record as the fold kernels
// do. A scene that converged sits out; the loop ends for everyone when no scene is running
// (the running count is read right after the first barrier, where no workgroup changes it) or
// after iteration cap. Partial sums alternate between two buffers, so a fold never reads a
// buffer that the next phase writes. Results are those of the chunked launches bit for bit.
template <bool RECOMP, bool CUBIC, bool COH>
__global__ void __launch_bounds__(kBlock) k_cg_persistent(Scenes sc, Nbr nb, Halo hl, const float2* __restrict__ cab,
                                                        const float* __restrict__ ca, const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel, const float* __restrict__ rho,
                                                        const float* __restrict__ irho, f3* __restrict__ r,
                                                        f3* __restrict__ pbuf0, f3* __restrict__ pbuf1, f3* __restrict__ Ap,
                                                        f3* __restrict__ x, SolverDev* stv, uint32_t* active,
                                                        double* __restrict__ partA, double* __restrict__ partB,
                                                        GridSync* gs, uint32_t cap, double tol, uint32_t spin) {
    const uint32_t nwg = gridDim.x;
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    const StepParams P = sc.P[sid];
    SolverDev* st = stv + sid;
    const bool writer = blk == P.blk0 && threadIdx.x == 0;
    __shared__ uint32_t running_s;
    bool done = st->done != 0u;
    double rr = st->rr;
    const double thr = st->thr;
    float beta = st->beta;
    for (uint32_t k = 1; k <= cap; k++) {
        f3* pprev = (k & 1u) ? pbuf0 : pbuf1;  // p of iteration k - 1 (p_0 = r in pbuf0)
        f3* pout = (k & 1u) ? pbuf1 : pbuf0;
        if (!done)
            cg_apply_block<1, RECOMP, CUBIC, COH>(blk, sid, beta, sc, nb, hl, cab, ca, pos, vel, rho, irho, nullptr, r, pprev,
                                                  pout, Ap, nullptr, nullptr, nullptr, partA);
        if (!grid_barrier<COH>(gs, nwg, spin)) return;
        if (threadIdx.x == 0) running_s = __hip_atomic_load(active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        if (running_s == 0u) return;
        float alpha = 0.0f;
        if (!done) {
            const double pAp = fold_as_fold_kernel<COH>(partA + P.blk0, P.nblk);
            alpha = float(rr / pAp);
            if (writer) {
                st->pAp = pAp;
                st->alpha = alpha;
            }
            cg_update3_block<COH>(blk, P.end, alpha, x, r, pout, Ap, partB);
        }
        if (!grid_barrier<COH>(gs, nwg, spin)) return;
        if (!done) {
            const double s = fold_as_fold_kernel<COH>(partB + P.blk0, P.nblk);
            beta = float(s / rr);
            rr = s;
            done = s < thr || (tol < 0.0 && s == 0.0);
            if (writer) {
                st->beta = beta;
                st->rr = s;
                if (k < MPS_MAX_ITER_LOG) st->log[k] = s;
                st->l = k + 1u;
                if (done) {
                    st->done = 1u;
                    atomicSub(active, 1u);
                }
            }
        }
    }
}

// ----------------------------------------------------------------- folds in the last workgroup
// The chunked CG without fold launches: each workgroup stores its partial device-coherently,
// waits until the store is performed and counts itself; the scene's last workgroup to arrive
// then sees every partial, folds them in the fold kernels' order (fold_as_fold_kernel, so the
// scalars have the same bits) and updates the record as k_fold_cg<MODE> would, for the next
// launch. Nothing waits for anything: the other workgroups simply end. Arrivals are counted in
// two levels, as same-address atomics serialize: groups of kArriveGroup consecutive workgroups
// of the scene count in arrive[blk0 + group], and the last of each group counts the group in
// the solver record's counter. Every counter is back at zero when the launch ends.
constexpr uint32_t kArriveGroup = 64;
template <int MODE>
__device__ __forceinline__ void fold_in_last_workgroup(const Scenes& sc, uint32_t sid, uint32_t blk,
                                                       const double* __restrict__ partials, SolverDev* stv, uint32_t* active,
                                                       double tol, uint32_t* arrive) {
    __shared__ uint32_t last_s;
    const StepParams P = sc.P[sid];
    SolverDev* st = stv + sid;
    __builtin_amdgcn_s_waitcnt(0);  // the partial (thread 0's coherent store) is performed
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t grp = (blk - P.blk0) / kArriveGroup;
        const uint32_t gsize = min(kArriveGroup, P.nblk - grp * kArriveGroup);
        const uint32_t ngrp = (P.nblk + kArriveGroup - 1u) / kArriveGroup;
        uint32_t* gc = arrive + P.blk0 + grp;
        bool last = false;
        if (__hip_atomic_fetch_add(gc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == gsize) {
            __hip_atomic_store(gc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = __hip_atomic_fetch_add(&st->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u == ngrp;
        }
        last_s = last;
    }
    __syncthreads();
    if (!last_s) return;
    const double s = fold_as_fold_kernel<true>(partials + P.blk0, P.nblk);
    if (threadIdx.x != 0) return;
    __hip_atomic_store(&st->counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (MODE == 1) {
        st->pAp = s;
        st->alpha = float(st->rr / s);
    } else {
        const uint32_t it = st->l;
        st->beta = float(s / st->rr);
        st->rr = s;
        if (it < MPS_MAX_ITER_LOG) st->log[it] = s;
        st->l = it + 1;
        if (s < st->thr || (tol < 0.0 && s == 0.0)) {
            st->done = 1u;
            atomicSub(active, 1u);
        }
    }
}
// k_cg_apply<1> + k_fold_cg<1> in one launch
template <bool RECOMP, bool CUBIC>
__global__ void __launch_bounds__(kBlock) k_cg_apply_fold(Scenes sc, Nbr nb, Halo hl, const float2* __restrict__ cab,
                                                        const float* __restrict__ ca, const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel, const float* __restrict__ rho,
                                                        const float* __restrict__ irho, const f3* __restrict__ rr,
                                                        const f3* __restrict__ pprev, f3* __restrict__ pout,
                                                        f3* __restrict__ Ap, SolverDev* st, double* __restrict__ partials,
                                                        uint32_t* arrive) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t sid = sc.blk_scene[blk];
    if (st[sid].done) return;
    cg_apply_block<1, RECOMP, CUBIC, false, true>(blk, sid, st[sid].beta, sc, nb, hl, cab, ca, pos, vel, rho, irho, nullptr, rr,
                                                  pprev, pout, Ap, nullptr, nullptr, nullptr, partials);
    fold_in_last_workgroup<1>(sc, sid, blk, partials, st, nullptr, 0.0, arrive);
}
// k_cg_update3 + k_fold_cg<2> in one launch
__global__ void __launch_bounds__(kBlock) k_cg_update3_fold(Scenes sc, f3* __restrict__ x, f3* __restrict__ r,
                                                          const f3* __restrict__ p, const f3* __restrict__ Ap, SolverDev* st,
                                                          double* __restrict__ partials, uint32_t* active, double tol,
                                                          uint32_t* arrive) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    if (st[sid].done) return;
    cg_update3_block<false, true>(blockIdx.x, sc.P[sid].end, st[sid].alpha, x, r, p, Ap, partials);
    fold_in_last_workgroup<2>(sc, sid, blockIdx.x, partials, st, active, tol, arrive);
}

// UpdateCGDirection_kernel (MPSCGKernel.cuh:123-133), p kept in the position records
__global__ void __launch_bounds__(kBlock) k_cg_direction(Scenes sc, const float4* __restrict__ r, float4* __restrict__ prec,
                                                       const SolverDev* st) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    if (st[sid].done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sc.P[sid].end) return;
    const float bta = st[sid].beta;
    const float4 ri = r[i];
    float4 a = prec[2 * size_t(i)], b = prec[2 * size_t(i) + 1];
    a.w = ri.x + bta * a.w;
    b.x = ri.y + bta * b.x;
    b.y = ri.z + bta * b.y;
    prec[2 * size_t(i)] = a;
    prec[2 * size_t(i) + 1] = b;
}

__global__ void __launch_bounds__(kBlock) k_copy_vel_out3(Scenes sc, const f3* __restrict__ x, float4* __restrict__ vel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sc.P[sc.blk_scene[blockIdx.x]].end) return;
    const f3 xi = x[i];
    float4 v = vel[i];
    v.x = xi.x;
    v.y = xi.y;
    v.z = xi.z;
    vel[i] = v;
}
__global__ void __launch_bounds__(kBlock) k_copy_vel_out(Scenes sc, const float4* __restrict__ x, float4* __restrict__ vel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= sc.P[sc.blk_scene[blockIdx.x]].end) return;
    const float4 xi = x[i];
    float4 v = vel[i];
    v.x = xi.x;
    v.y = xi.y;
    v.z = xi.z;
    vel[i] = v;
}

// ----------------------------------------------------------------- Jacobi viscosity
// ComputeJacobiViscosity kernels (MPSSPHNonPressureUtil.cuh:90-194, commented out in the
// reference): per particle a 3x3 block with A_ij = V G/((d^2+0.01h^2)(d+eps)) x_ij x_ij^T.
template <bool CUBIC>
__global__ void __launch_bounds__(kBlock) k_jacobi_candidate(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                           const float4* __restrict__ vel, const float* __restrict__ rho,
                                                           const float4* __restrict__ u, float4* __restrict__ out) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    if (i >= P.end) return;
    const float4 xi4 = pos[i];
    const f3 xi = mk(xi4);
    const float h = xi4.w * 4.0f;
    const float invH = 1.0f / h;
    const float onePercentHSq = 0.01f * h * h;
    float A[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    f3 Avj{0.f, 0.f, 0.f};
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const uint32_t j = nbr_ff(nb, k, i, N);
        const f3 xij = xi - mk(pos[j]);
        const float dist = len3(xij);
        if (dist < h) {
            const f3 uj = mk(u[j]);
            const float cc = P.V * Gk<CUBIC>(dist, invH) / ((dist * dist + onePercentHSq) * (dist + kFltEps));
            const float M[3][3] = {{cc * (xij.x * xij.x), cc * (xij.y * xij.x), cc * (xij.z * xij.x)},
                                   {cc * (xij.x * xij.y), cc * (xij.y * xij.y), cc * (xij.z * xij.y)},
                                   {cc * (xij.x * xij.z), cc * (xij.y * xij.z), cc * (xij.z * xij.z)}};
            const f3 uM{dot3(uj, f3{M[0][0], M[0][1], M[0][2]}), dot3(uj, f3{M[1][0], M[1][1], M[1][2]}),
                        dot3(uj, f3{M[2][0], M[2][1], M[2][2]})};
            Avj = Avj - uM;
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int bq = 0; bq < 3; bq++) A[a][bq] = A[a][bq] - M[a][bq];
        }
    }
    const float al = P.dt * 10.0f * P.mu / rho[i];
    Avj = Avj * al;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int bq = 0; bq < 3; bq++) A[a][bq] = A[a][bq] * al + (a == bq ? 1.0f : 0.0f);
    const float oneOverDet = 1.0f / (+A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]) -
                                     A[1][0] * (A[0][1] * A[2][2] - A[2][1] * A[0][2]) +
                                     A[2][0] * (A[0][1] * A[1][2] - A[1][1] * A[0][2]));
    float I[3][3];
    I[0][0] = +(A[1][1] * A[2][2] - A[2][1] * A[1][2]) * oneOverDet;
    I[1][0] = -(A[1][0] * A[2][2] - A[2][0] * A[1][2]) * oneOverDet;
    I[2][0] = +(A[1][0] * A[2][1] - A[2][0] * A[1][1]) * oneOverDet;
    I[0][1] = -(A[0][1] * A[2][2] - A[2][1] * A[0][2]) * oneOverDet;
    I[1][1] = +(A[0][0] * A[2][2] - A[2][0] * A[0][2]) * oneOverDet;
    I[2][1] = -(A[0][0] * A[2][1] - A[2][0] * A[0][1]) * oneOverDet;
    I[0][2] = +(A[0][1] * A[1][2] - A[1][1] * A[0][2]) * oneOverDet;
    I[1][2] = -(A[0][0] * A[1][2] - A[1][0] * A[0][2]) * oneOverDet;
    I[2][2] = +(A[0][0] * A[1][1] - A[1][0] * A[0][1]) * oneOverDet;
    if (isinf(I[0][0]) || isnan(I[0][0])) {
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int bq = 0; bq < 3; bq++) I[a][bq] = 0.0f;
    }
    const f3 rhs = mk(vel[i]) + Avj;
    out[i] = make_float4(dot3(rhs, f3{I[0][0], I[0][1], I[0][2]}), dot3(rhs, f3{I[1][0], I[1][1], I[1][2]}),
                         dot3(rhs, f3{I[2][0], I[2][1], I[2][2]}), 0.0f);
}

__global__ void __launch_bounds__(kBlock) k_jacobi_relax(uint32_t N, const float4* __restrict__ cand, const float4* __restrict__ cur,
                                                       const float4* __restrict__ prev, float ur, float w0, float w1, float w2,
                                                       float4* __restrict__ o0, float4* __restrict__ o1, float4* __restrict__ o2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const f3 nw = mk(cand[i]), c = mk(cur[i]), pv = mk(prev[i]);
    const f3 t = ur * (nw - c) + c - pv;
    const f3 a = w0 * t + pv, b = w1 * t + pv, d = w2 * t + pv;
    o0[i] = make_float4(a.x, a.y, a.z, 0.f);
    o1[i] = make_float4(b.x, b.y, b.z, 0.f);
    o2[i] = make_float4(d.x, d.y, d.z, 0.f);
}

// ComputeJacobiError_0/1_kernel (MPSSPHNonPressureUtil.cuh:195-333): per-block partial sums
// of |u_i - a_i sum_j (x_ij.(u_i-u_j)) grad_ij - v0_i| for up to three candidates.
template <bool CUBIC>
__global__ void __launch_bounds__(kBlock) k_jacobi_error(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                       const float4* __restrict__ vel, const float* __restrict__ rho,
                                                       const float4* __restrict__ u0, const float4* __restrict__ u1,
                                                       const float4* __restrict__ u2, int ncand, double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    const float4* U[3] = {u0, u1, u2};
    for (int q = 0; q < ncand; q++) {
        double e = 0.0;
        if (i < P.end) {
            const float4 xi4 = pos[i];
            const f3 xi = mk(xi4);
            const float h = xi4.w * 4.0f;
            const float invH = 1.0f / h;
            const float onePercentHSq = 0.01f * h * h;
            const f3 ui = mk(U[q][i]);
            f3 Av{0.f, 0.f, 0.f};
            const uint32_t c = nb.cff[i];
            for (uint32_t k = 0; k < c; k++) {
                const uint32_t j = nbr_ff(nb, k, i, N);
                const f3 xij = xi - mk(pos[j]);
                const float dist = len3(xij);
                if (dist < h) {
                    const f3 grad = P.V * Gk<CUBIC>(dist, invH) / ((dist * dist + onePercentHSq) * (dist + kFltEps)) * xij;
                    Av = Av + dot3(xij, ui - mk(U[q][j])) * grad;
                }
            }
            Av = Av * (P.dt * 10.0f * P.mu / rho[i]);
            e = double(len3(ui - Av - mk(vel[i])));
        }
        const double bs = block_sum_d(e, lds);
        if (threadIdx.x == 0) partials[size_t(q) * gridDim.x + blockIdx.x] = bs;
    }
}
__global__ void __launch_bounds__(kBlock) k_fold_partials(const double* __restrict__ partials, uint32_t nblk, int ncand,
                                                        double* __restrict__ out) {
    __shared__ double lds[kBlock / 64];
    for (int q = 0; q < ncand; q++) {
        double s = 0.0;
        for (uint32_t k = threadIdx.x; k < nblk; k += blockDim.x) s += partials[size_t(q) * nblk + k];
        s = block_sum_d(s, lds);
        if (threadIdx.x == 0) out[q] = s;
    }
}

// ----------------------------------------------------------------- advection
// BoundaryCollision_kernel + UpdatePosition_kernel (MPSAdvectUtil.cuh:36-77)
__global__ void __launch_bounds__(kBlock) k_advect(Scenes sc, float4* __restrict__ pos, float4* __restrict__ vel,
                                                 uint32_t* __restrict__ nonfinite) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.end) return;
    float4 x4 = pos[i];
    float4 v4 = vel[i];
    float v[3] = {v4.x, v4.y, v4.z};
    const float x[3] = {x4.x, x4.y, x4.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float pk = x[k] + v[k] * P.dt;
        if (pk < P.bmin[k] && v[k] < 0.0f) v[k] = (fabsf(v[k]) < 1.0e-4f) ? 0.0f : -v[k] * 0.4f;
        else if (pk > P.bmax[k] && v[k] > 0.0f) v[k] = (fabsf(v[k]) < 1.0e-4f) ? 0.0f : -v[k] * 0.4f;
    }
    v4.x = v[0];
    v4.y = v[1];
    v4.z = v[2];
    x4.x = x[0] + P.dt * v[0];
    x4.y = x[1] + P.dt * v[1];
    x4.z = x[2] + P.dt * v[2];
    vel[i] = v4;
    pos[i] = x4;
    const bool bad = !isfinite(x4.x) || !isfinite(x4.y) || !isfinite(x4.z) || !isfinite(v4.x) || !isfinite(v4.y) ||
                     !isfinite(v4.z);
    if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(nonfinite, 1u);
}

// One workgroup per scene: st[blockIdx.x] starts a new loop; *active = S.
__global__ void k_solver_reset(SolverDev* stv, uint32_t S, uint32_t* active) {
    SolverDev* st = stv + blockIdx.x;
    const uint32_t t = threadIdx.x;
    if (t == 0) {
        st->done = 0u;
        st->l = 1u;
        st->counter = 0u;
        st->rr = 0.0;
        st->pAp = 0.0;
        st->thr = 0.0;
        st->alpha = 0.0f;
        st->beta = 0.0f;
        if (blockIdx.x == 0) *active = S;
    }
    for (uint32_t k = t; k < MPS_MAX_ITER_LOG; k += blockDim.x) st->log[k] = 0.0;
}
// Plain streaming copy (16-byte vectors), used only to calibrate byte counters.
__global__ void __launch_bounds__(kBlock) k_calib_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) b[i] = a[i];
}
// STREAM-style copy for measuring what the memory system gives a copy: one 16-byte vector per
// thread (no grid-stride loop), non-temporal loads and stores. Measured on MI355X: 6.7 TB/s
// against 4.8 TB/s for the grid-stride k_calib_copy above, which stays as the byte-counter
// reference (its loads and stores are ordinary ones, counted like the kernels' own).
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(kBlock) k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n)
        __builtin_nontemporal_store(__builtin_nontemporal_load(reinterpret_cast<const v4f*>(a) + i), reinterpret_cast<v4f*>(b) + i);
}
__global__ void k_zero_words(uint32_t* __restrict__ p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
// Render feed: positions and radii de-interleaved for the host (replaces the GL buffers
// SPHRenderer binds as SSBOs, MPSSPHRenderer.cpp:22-29).
__global__ void __launch_bounds__(kBlock) k_pack_readback(uint32_t n, const float4* __restrict__ pos, const float4* __restrict__ color,
                                                        const uint32_t* __restrict__ perm, float* __restrict__ xyz,
                                                        float* __restrict__ rad, float* __restrict__ rgba) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm[s];  // reference index of slot s
    if (i == kEmpty) return;     // padding between scenes
    const float4 p = pos[s];
    if (xyz) {
        xyz[3 * size_t(i)] = p.x;
        xyz[3 * size_t(i) + 1] = p.y;
        xyz[3 * size_t(i) + 2] = p.z;
    }
    if (rad) rad[i] = p.w;
    if (rgba) reinterpret_cast<float4*>(rgba)[i] = color[s];
}

// The same for a renderer that reads double positions and radii (SPHRenderer's shader,
// src/shader/particleShader.vert:3-4): the float values widened exactly.
__global__ void __launch_bounds__(kBlock) k_pack_readback_f64(uint32_t n, const float4* __restrict__ pos,
                                                            const float4* __restrict__ color, const uint32_t* __restrict__ perm,
                                                            double* __restrict__ xyz, double* __restrict__ rad,
                                                            float* __restrict__ rgba) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t i = perm[s];
    if (i == kEmpty) return;
    const float4 p = pos[s];
    if (xyz) {
        xyz[3 * size_t(i)] = double(p.x);
        xyz[3 * size_t(i) + 1] = double(p.y);
        xyz[3 * size_t(i) + 2] = double(p.z);
    }
    if (rad) rad[i] = double(p.w);
    if (rgba) reinterpret_cast<float4*>(rgba)[i] = color[s];
}

}  // namespace dev

// ===================================================================== host launchers
namespace launch {
using namespace dev;
static inline unsigned nb_(size_t n) { return unsigned((n + kBlock - 1) / kBlock); }

void boundary_sum(int cubic, const float4* bpos, uint32_t B, const uint32_t* nbb, uint32_t Kbb, const uint32_t* cbb, float* S,
                  hipStream_t s) {
    if (!B) return;
    if (cubic) launch_k(k_boundary_sum<true>, nb_(B), kBlock, 0, s, bpos, B, nbb, Kbb, cbb, S);
    else launch_k(k_boundary_sum<false>, nb_(B), kBlock, 0, s, bpos, B, nbb, Kbb, cbb, S);
}
void boundary_sum_obj(int cubic, const float4* bpos, uint32_t B, const uint32_t* nbb, uint32_t Kbb, const uint32_t* cbb,
                      const uint32_t* obj, float4* S4, hipStream_t s) {
    if (!B) return;
    launch_k(cubic ? k_boundary_sum_obj<true> : k_boundary_sum_obj<false>, nb_(B), kBlock, 0, s, bpos, B, nbb, Kbb, cbb, obj, S4);
}
void boundary_volume_obj(const float4* S4, const uint32_t* obj, uint32_t nobj, float* vol, uint32_t B, int faithful,
                         hipStream_t s) {
    if (B) launch_k(k_boundary_volume_obj, nb_(B), kBlock, 0, s, S4, obj, nobj, vol, B, faithful);
}
void boundary_volume(const float* S, float* vol, uint32_t B, int faithful, hipStream_t s) {
    if (B) launch_k(k_boundary_volume, nb_(B), kBlock, 0, s, S, vol, B, faithful);
}
void density_alpha(int cubic, int flags, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, float4* vel,
                   const float4* bpos, const float* bvol, float* rho, float* alpha, float4* color, float* pG_ff,
                   float* pG_fb, hipStream_t s) {
    if (!sc.N) return;
    using K = decltype(&k_density_alpha<false, false, false>);
    static constexpr K tab[8] = {k_density_alpha<false, false, false>, k_density_alpha<false, false, true>,
                                 k_density_alpha<false, true, false>,  k_density_alpha<false, true, true>,
                                 k_density_alpha<true, false, false>,  k_density_alpha<true, false, true>,
                                 k_density_alpha<true, true, false>,   k_density_alpha<true, true, true>};
    const K k = tab[(cubic ? 4 : 0) | (nb.fr ? 2 : 0) | (nb.fbo ? 1 : 0)];
    launch_k(k, nb_(sc.N), kBlock, 0, s, flags, sc, nb, hl, pos, vel, bpos, bvol, rho, alpha, color, pG_ff, pG_fb);
}
void pressure_eval(int cubic, int df, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const float4* vel,
                   const float4* bpos, const float* bvol, const float4* bvel, const float* rho, const float* alpha,
                   const float* pG_ff, const float* pG_fb, float* kappa, float* delta, const SolverDev* st, double* partials,
                   hipStream_t s) {
    if (!sc.N) return;
    const unsigned g = nb_(sc.N);
    // the ref sums evaluate G_ij themselves (the stored G covers the own and wall lists only);
    // `cubic` matters for the ref instances only
    using K = decltype(&k_pressure_eval<true, false, false, false>);
    static constexpr K tab[16] = {
        k_pressure_eval<false, false, false, false>, k_pressure_eval<false, false, false, true>,
        k_pressure_eval<false, false, false, false>, k_pressure_eval<false, false, false, true>,
        k_pressure_eval<false, true, false, false>,  k_pressure_eval<false, true, false, true>,
        k_pressure_eval<false, true, true, false>,   k_pressure_eval<false, true, true, true>,
        k_pressure_eval<true, false, false, false>,  k_pressure_eval<true, false, false, true>,
        k_pressure_eval<true, false, false, false>,  k_pressure_eval<true, false, false, true>,
        k_pressure_eval<true, true, false, false>,   k_pressure_eval<true, true, false, true>,
        k_pressure_eval<true, true, true, false>,    k_pressure_eval<true, true, true, true>};
    const K k = tab[(df ? 8 : 0) | (nb.fr ? 4 : 0) | (cubic ? 2 : 0) | (nb.fbo ? 1 : 0)];
    launch_k(k, g, kBlock, 0, s, sc, nb, hl, pos, vel, bpos, bvol, bvel, rho, alpha, pG_ff, pG_fb, kappa, delta, st, partials);
}
void fold_pressure(const Scenes& sc, const double* partials, SolverDev* st, uint32_t* active, double tol, uint32_t lmin,
                   int joint, hipStream_t s) {
    if (!sc.N) return;
    if (joint) launch_k(k_fold_pressure_joint, 1, kFoldThreads, 0, s, sc, partials, st, active, tol, lmin);
    else launch_k(k_fold_pressure, sc.S, kFoldThreads, 0, s, sc, partials, st, active, tol, lmin);
}
void pressure_apply(int cubic, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, float4* vel, const float4* bpos,
                    const float* bvol, const float* pG_ff, const float* pG_fb, const float* kappa, const SolverDev* st,
                    hipStream_t s) {
    if (!sc.N) return;
    using K = decltype(&k_pressure_apply<false, false, false>);
    static constexpr K tab[8] = {k_pressure_apply<false, false, false>, k_pressure_apply<false, false, true>,
                                 k_pressure_apply<false, false, false>, k_pressure_apply<false, false, true>,
                                 k_pressure_apply<true, false, false>,  k_pressure_apply<true, false, true>,
                                 k_pressure_apply<true, true, false>,   k_pressure_apply<true, true, true>};
    const K k = tab[(nb.fr ? 4 : 0) | (cubic ? 2 : 0) | (nb.fbo ? 1 : 0)];
    launch_k(k, nb_(sc.N), kBlock, 0, s, sc, nb, hl, pos, vel, bpos, bvol, pG_ff, pG_fb, kappa, st);
}
void st_factor(const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const float4* vel, const float* pG_ff, float* fst,
               float4* cgx, hipStream_t s) {
    if (sc.N) launch_k(k_st_factor, nb_(sc.N), kBlock, 0, s, sc, nb, hl, pos, vel, pG_ff, fst, cgx);
}
void cg_setup(int cubic, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const float4* vel, const float* rho,
              const float* fst, float* ca, float2* cab, float4* cgb, float* irho, hipStream_t s) {
    if (!sc.N) return;
    if (cubic) launch_k(k_cg_setup<true>, nb_(sc.N), kBlock, 0, s, sc, nb, hl, pos, vel, rho, fst, ca, cab, cgb, irho);
    else launch_k(k_cg_setup<false>, nb_(sc.N), kBlock, 0, s, sc, nb, hl, pos, vel, rho, fst, ca, cab, cgb, irho);
}
void cg_pack(uint32_t N, const float4* pos, const float4* x0, float4* rec, hipStream_t s) {
    if (N) launch_k(k_cg_pack, nb_(N), kBlock, 0, s, N, pos, x0, rec);
}
void cg_apply(int mode, const CgOperator& op, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos,
              const float4* x0, const f3* rr, const f3* pprev, f3* pout, f3* Ap, const float4* b, f3* r, f3* x,
              const SolverDev* st, double* partials, hipStream_t s) {
    if (!sc.N) return;
    const unsigned g = nb_(sc.N);
#define MPS_CG_APPLY(M, R, C, ...)                                                                                   \
    launch_k(k_cg_apply<M, R, C, ##__VA_ARGS__>, g, kBlock, 0, s, sc, nb, hl, op.cab, op.ca, pos, op.vel, op.rho, op.irho, \
             x0, rr, pprev, pout, Ap, b, r, x, st, partials)
    if (!op.recompute && op.deep) {
        if (mode == 0) MPS_CG_APPLY(0, false, true, true);
        else MPS_CG_APPLY(1, false, true, true);
    } else if (!op.recompute) {
        if (mode == 0) MPS_CG_APPLY(0, false, true);
        else MPS_CG_APPLY(1, false, true);
    } else if (op.cubic) {
        if (mode == 0) MPS_CG_APPLY(0, true, true);
        else MPS_CG_APPLY(1, true, true);
    } else {
        if (mode == 0) MPS_CG_APPLY(0, true, false);
        else MPS_CG_APPLY(1, true, false);
    }
#undef MPS_CG_APPLY
}
void cg_apply_fold(const CgOperator& op, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const f3* rr,
                   const f3* pprev, f3* pout, f3* Ap, SolverDev* st, double* partials, uint32_t* arrive, hipStream_t s) {
    if (!sc.N) return;
    const unsigned g = nb_(sc.N);
#define MPS_CG_APPLY_FOLD(R, C)                                                                                    \
    launch_k(k_cg_apply_fold<R, C>, g, kBlock, 0, s, sc, nb, hl, op.cab, op.ca, pos, op.vel, op.rho, op.irho, rr, pprev, \
             pout, Ap, st, partials, arrive)
    if (!op.recompute) MPS_CG_APPLY_FOLD(false, true);
    else if (op.cubic) MPS_CG_APPLY_FOLD(true, true);
    else MPS_CG_APPLY_FOLD(true, false);
#undef MPS_CG_APPLY_FOLD
}
void cg_update3_fold(const Scenes& sc, f3* x, f3* r, const f3* p, const f3* Ap, SolverDev* st, double* partials,
                     uint32_t* active, double tol, uint32_t* arrive, hipStream_t s) {
    if (sc.N) launch_k(k_cg_update3_fold, nb_(sc.N), kBlock, 0, s, sc, x, r, p, Ap, st, partials, active, tol, arrive);
}
static const void* persistent_fn(const CgOperator& op, int coherent) {
    if (coherent) {
        if (!op.recompute) return reinterpret_cast<const void*>(k_cg_persistent<false, true, true>);
        return op.cubic ? reinterpret_cast<const void*>(k_cg_persistent<true, true, true>)
                        : reinterpret_cast<const void*>(k_cg_persistent<true, false, true>);
    }
    if (!op.recompute) return reinterpret_cast<const void*>(k_cg_persistent<false, true, false>);
    return op.cubic ? reinterpret_cast<const void*>(k_cg_persistent<true, true, false>)
                    : reinterpret_cast<const void*>(k_cg_persistent<true, false, false>);
}
uint32_t cg_persistent_capacity(const CgOperator& op, int coherent) {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, persistent_fn(op, coherent), kBlock, 0) != hipSuccess) return 0;
    return uint32_t(per_cu) * uint32_t(prop.multiProcessorCount);
}
void cg_persistent(const CgOperator& op, int coherent, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos,
                   f3* r, f3* pbuf0, f3* pbuf1, f3* Ap, f3* x, SolverDev* st, uint32_t* active, double* partA,
                   double* partB, GridSync* gs, uint32_t cap, double tol, uint32_t spin, hipStream_t s) {
    if (!sc.N) return;
    const uint32_t nblk = nb_(sc.N);
    void* args[] = {const_cast<Scenes*>(&sc), const_cast<Nbr*>(&nb), const_cast<Halo*>(&hl), const_cast<float2**>(&op.cab),
                    const_cast<float**>(&op.ca), const_cast<float4**>(&pos), const_cast<float4**>(&op.vel),
                    const_cast<float**>(&op.rho), const_cast<float**>(&op.irho), &r, &pbuf0, &pbuf1, &Ap, &x, &st, &active,
                    &partA, &partB, &gs, &cap, &tol, &spin};
    const hipError_t rc = hipLaunchCooperativeKernel(persistent_fn(op, coherent), dim3(nblk), dim3(kBlock), args, 0, s);
    if (rc != hipSuccess) throw LaunchError{rc};
}
void fold_cg(int mode, const Scenes& sc, const double* partials, SolverDev* st, uint32_t* active, double tol, hipStream_t s) {
    if (!sc.N) return;
    if (mode == 0) launch_k(k_fold_cg<0>, sc.S, kFoldThreads, 0, s, sc, partials, st, active, tol);
    else if (mode == 1) launch_k(k_fold_cg<1>, sc.S, kFoldThreads, 0, s, sc, partials, st, active, tol);
    else launch_k(k_fold_cg<2>, sc.S, kFoldThreads, 0, s, sc, partials, st, active, tol);
}
void cg_update(const Scenes& sc, float4* x, float4* r, const float4* prec, const float4* Ap, const SolverDev* st,
               double* partials, hipStream_t s) {
    if (sc.N) launch_k(k_cg_update, nb_(sc.N), kBlock, 0, s, sc, x, r, prec, Ap, st, partials);
}
void cg_update3(const Scenes& sc, f3* x, f3* r, const f3* p, const f3* Ap, const SolverDev* st, double* partials,
                hipStream_t s) {
    if (sc.N) launch_k(k_cg_update3, nb_(sc.N), kBlock, 0, s, sc, x, r, p, Ap, st, partials);
}
void copy_vel_out3(const Scenes& sc, const f3* x, float4* vel, hipStream_t s) {
    if (sc.N) launch_k(k_copy_vel_out3, nb_(sc.N), kBlock, 0, s, sc, x, vel);
}
void cg_direction(const Scenes& sc, const float4* r, float4* prec, const SolverDev* st, hipStream_t s) {
    if (sc.N) launch_k(k_cg_direction, nb_(sc.N), kBlock, 0, s, sc, r, prec, st);
}
void copy_vel_out(const Scenes& sc, const float4* x, float4* vel, hipStream_t s) {
    if (sc.N) launch_k(k_copy_vel_out, nb_(sc.N), kBlock, 0, s, sc, x, vel);
}
void jacobi_candidate(int cubic, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel,
                      const float* rho, const float4* u, float4* out, hipStream_t s) {
    if (!sc.N) return;
    if (cubic) launch_k(k_jacobi_candidate<true>, nb_(sc.N), kBlock, 0, s, sc, nb, pos, vel, rho, u, out);
    else launch_k(k_jacobi_candidate<false>, nb_(sc.N), kBlock, 0, s, sc, nb, pos, vel, rho, u, out);
}
void jacobi_relax(uint32_t N, const float4* cand, const float4* cur, const float4* prev, float ur, const float* w3,
                  float4* o0, float4* o1, float4* o2, hipStream_t s) {
    if (N) launch_k(k_jacobi_relax, nb_(N), kBlock, 0, s, N, cand, cur, prev, ur, w3[0], w3[1], w3[2], o0, o1, o2);
}
void jacobi_error(int cubic, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel,
                  const float* rho, const float4* u0, const float4* u1, const float4* u2, int ncand, double* partials,
                  double* err3, hipStream_t s) {
    if (!sc.N) return;
    const unsigned g = nb_(sc.N);
    if (cubic) launch_k(k_jacobi_error<true>, g, kBlock, 0, s, sc, nb, pos, vel, rho, u0, u1, u2, ncand, partials);
    else launch_k(k_jacobi_error<false>, g, kBlock, 0, s, sc, nb, pos, vel, rho, u0, u1, u2, ncand, partials);
    launch_k(k_fold_partials, 1, kBlock, 0, s, partials, g, ncand, err3);
}
void advect(const Scenes& sc, float4* pos, float4* vel, uint32_t* nonfinite, hipStream_t s) {
    if (sc.N) launch_k(k_advect, nb_(sc.N), kBlock, 0, s, sc, pos, vel, nonfinite);
}
void solver_reset(SolverDev* st, uint32_t S, uint32_t* active, hipStream_t s) {
    if (S) launch_k(k_solver_reset, S, kBlock, 0, s, st, S, active);
}
void calib_copy(const float4* a, float4* b, size_t n, hipStream_t s) {
    if (n) launch_k(k_calib_copy, 4096, kBlock, 0, s, a, b, n);
}
void stream_copy(const float4* a, float4* b, size_t n, hipStream_t s) {
    if (n) launch_k(k_stream_copy, unsigned((n + kBlock - 1) / kBlock), kBlock, 0, s, a, b, n);
}
void zero_words(uint32_t* p, size_t n, hipStream_t s) {
    if (n) launch_k(k_zero_words, unsigned((n + kBlock - 1) / kBlock), kBlock, 0, s, p, n);
}
void pack_readback(uint32_t n, const float4* pos, const float4* color, const uint32_t* perm, float* xyz, float* rad,
                   float* rgba, hipStream_t s) {
    if (n) launch_k(k_pack_readback, nb_(n), kBlock, 0, s, n, pos, color, perm, xyz, rad, rgba);
}
void pack_readback_f64(uint32_t n, const float4* pos, const float4* color, const uint32_t* perm, double* xyz, double* rad,
                       float* rgba, hipStream_t s) {
    if (n) launch_k(k_pack_readback_f64, nb_(n), kBlock, 0, s, n, pos, color, perm, xyz, rad, rgba);
}

}  // namespace launch
}  // namespace mps
