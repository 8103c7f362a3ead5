// adjoint.hip — reverse-mode kernels of the SPH step for gfx950.
//
// The differentiated map is one step with its control flow frozen: the step's neighbour
// lists, the iteration counts of the DF and CD loops, and the branch every particle took (the
// positive-error mask and the min of the DF stiffness, the collision response). The DF/CD
// loops are differentiated through their iterations. The implicit viscosity / surface-tension
// solve is differentiated with the implicit-function theorem: the solve converges to
//   R(v*) = A_f v* - b - f v_in = 0
// (its residual starts from A with f = 0 and iterates with A_f), so with A_f^T lambda = vbar*
// every input theta gets thetabar = -lambda . dR/dtheta.
//
// Kernel derivatives in closed form for both kernel sets (cubic spline; poly6 density with the
// spiky gradient): W' = WG d (WG = G for the cubic spline), STW' = STG d, G' and STG';
// J_ij = d(G_ij x_ij)/dx_i = G I + (G'/d) x_ij x_ij^T. The neighbour test |x_ij| < (h_i + h_j)/2
// is symmetric, so the lists are symmetric and every pair term is gathered at both of its ends
// (no kernel needs atomics). Each end evaluates the kernel with its own support h = 4 r: with
// equal radii in every scene (EQ) G_ji = G_ij and the stored value serves both ends; otherwise the
// neighbour's G_ji and G'_ji are recomputed from h_j.
#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mps {
namespace dev {

__device__ __forceinline__ f3 f3z() { return {0.f, 0.f, 0.f}; }
__device__ __forceinline__ f3 neg(f3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float4 f4(f3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// dG/dd: cubic spline (G = kCG invh^5 (3q - 2) for q <= 1/2, -kCG invh^4 (1-q)^2 / d above) or
// spiky (G = -kG invh^4 (1-q)^2 / d on (0, 1))
template <bool CUBIC>
__device__ __forceinline__ float Gpk(float d, float invh) {
    const float q = d * invh;
    if (q < kEps || q >= 1.0f) return 0.0f;
    const float h2 = invh * invh, h4 = h2 * h2;
    const float t = 1.0f - q;
    if (!CUBIC) return kG * h4 * t * (2.0f * invh / d + t / (d * d));
    if (q > 0.5f) return kCubicG * h4 * t * (2.0f * invh / d + t / (d * d));
    return 3.0f * kCubicG * h4 * h2;
}
// (dW/dd) / d of the density kernel: the cubic spline's G, or for poly6
// (W = kW ((1 - q^2) invh)^3) -6 kW invh^3 ((1 - q^2) invh)^2
template <bool CUBIC>
__device__ __forceinline__ float WGk(float d, float invh) {
    if (CUBIC) return Gk<true>(d, invh);
    const float q = d * invh;
    if (q >= 1.0f) return 0.0f;
    const float tmp = (1.0f - q * q) * invh;
    return -6.0f * kW * (invh * invh * invh) * (tmp * tmp);
}
// G and G' of the pair seen from the neighbour j: the values at i when the radii are equal
template <bool CUBIC, bool EQ>
__device__ __forceinline__ float2 other_end(float G, float Gp, float d, float4 xj4) {
    if (EQ) return make_float2(G, Gp);
    const float invHj = 1.0f / (xj4.w * 4.0f);
    return make_float2(Gk<CUBIC>(d, invHj), Gpk<CUBIC>(d, invHj));
}
// dSTG/dd (STG = -kSTG invh^4 (1-q)^2 / d on (1/2, 1))
__device__ __forceinline__ float STGp(float d, float invh) {
    const float q = d * invh;
    if (q <= 0.5f || q >= 1.0f) return 0.0f;
    const float h2 = invh * invh;
    const float t = 1.0f - q;
    return kSTG * h2 * h2 * t * (2.0f * invh / d + t / (d * d));
}
// J w with J = G I + (G'/d) x x^T
__device__ __forceinline__ f3 Jmul(float G, float Gp, float d, f3 x, f3 w) {
    const float c = d > 0.0f ? Gp / d : 0.0f;
    return G * w + (c * dot3(x, w)) * x;
}

// Workgroup partial of two doubles (fixed order), published at partials[blk] and
// partials[nblk + blk].
__device__ __forceinline__ void block_partial2(double a, double b, double* __restrict__ partials, uint32_t nblk) {
    __shared__ double lds[kBlock / 64];
    const double sa = block_sum_d(a, lds);
    const double sb = block_sum_d(b, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sa;
        partials[nblk + blockIdx.x] = sb;
    }
}

// ----------------------------------------------------------------- advection
// x' = x + dt v', v' = c (.) v with c in {1, -0.4, 0} per component (the collision branch of
// MPSAdvectUtil.cuh:36-77, recomputed from the same inputs).
__global__ void __launch_bounds__(kBlock) k_adj_advect(Scenes sc, const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                     const float4* __restrict__ ax, float4* __restrict__ av) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.end) return;
    const float4 x4 = pos[i], v4 = vin[i], a4 = ax[i], b4 = av[i];
    const float x[3] = {x4.x, x4.y, x4.z}, v[3] = {v4.x, v4.y, v4.z};
    const float xa[3] = {a4.x, a4.y, a4.z}, va[3] = {b4.x, b4.y, b4.z};
    float o[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float pk = x[k] + v[k] * P.dt;
        float c = 1.0f;
        if ((pk < P.bmin[k] && v[k] < 0.0f) || (!(pk < P.bmin[k] && v[k] < 0.0f) && pk > P.bmax[k] && v[k] > 0.0f))
            c = (fabsf(v[k]) < 1.0e-4f) ? 0.0f : -0.4f;
        o[k] = c * (va[k] + P.dt * xa[k]);
    }
    av[i] = make_float4(o[0], o[1], o[2], 0.0f);
}

// ----------------------------------------------------------------- DF / CD iterations
// Recompute delta, the stiffness kappa and the branch of every particle from the velocity the
// iteration started from (k_pressure_eval's arithmetic). aux = (kappa, e, mask, rho branch).
template <bool DF>
__global__ void __launch_bounds__(kBlock) k_adj_peval(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                    const float4* __restrict__ vel, const float4* __restrict__ bpos,
                                                    const float* __restrict__ bvol, const float4* __restrict__ bvel,
                                                    const float* __restrict__ rho, const float* __restrict__ alpha,
                                                    const float* __restrict__ pG_ff, const float* __restrict__ pG_fb,
                                                    float4* __restrict__ aux) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    if (i >= P.end) return;
    const f3 xi = mk(pos[i]), vi = mk(vel[i]);
    float df = 0.0f;
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.ff[slot];
        df += P.V * pG_ff[slot] * dot3(xi - mk(pos[j]), vi - mk(vel[j]));
    }
    float db = 0.0f;
    const uint32_t cb = nb.cfb[i];
    for (uint32_t k = 0; k < cb; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.fb[slot];
        db += bvol[j] * pG_fb[slot] * dot3(xi - mk(bpos[j]), vi - mk(bvel[j]));
    }
    float d = 0.0f;
    d += df;
    d += db;
    float e, rb;
    if (DF) {
        const float a = d * P.dt;
        const float b = rho[i] / P.rho0 + P.dt * d - 0.8f;
        e = a < b ? a : b;
        rb = a < b ? 0.0f : 1.0f;
    } else {
        e = rho[i] / P.rho0 + P.dt * d - 1.0f;
        rb = 1.0f;
    }
    float kap = 0.0f, m = 0.0f;
    if (e > 0.0f) {
        kap = e * (alpha[i] / (P.dt * P.dt));
        m = 1.0f;
    } else {
        e = 0.0f;
    }
    aux[i] = make_float4(kap, e, m, rb);
}

// kappabar from vbar of the velocity after the force (v_i += -dt/2 [sum_j V (k_i+k_j) G x +
// sum_b V_b k_i G x]), then deltabar, alphabar, rhobar, and the position terms of the force.
// Runs for scenes whose applied iterations include `it` (it < L[scene]).
template <bool DF, bool CUBIC, bool EQ>
__global__ void __launch_bounds__(kBlock) k_adj_kbar(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                   const float4* __restrict__ bpos, const float* __restrict__ bvol,
                                                   const float* __restrict__ pG_ff, const float* __restrict__ pG_fb,
                                                   const float4* __restrict__ aux, const float* __restrict__ alpha,
                                                   const float4* __restrict__ av, const uint32_t* __restrict__ L, uint32_t it,
                                                   float* __restrict__ dbar, float4* __restrict__ ax,
                                                   float* __restrict__ abar, float* __restrict__ rbar) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    if (it >= L[sid]) return;
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    if (i >= P.end) return;
    const float4 xi4 = pos[i];
    const f3 xi = mk(xi4);
    const float invH = 1.0f / (xi4.w * 4.0f);
    const f3 vbi = mk(av[i]);
    const float4 ai = aux[i];
    float kb = 0.0f;
    f3 gx = f3z();
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.ff[slot];
        const float4 xj4 = pos[j];
        const f3 xij = xi - mk(xj4);
        const float G = pG_ff[slot];
        const float d = len3(xij);
        const float Gp = Gpk<CUBIC>(d, invH);
        const float kk = ai.x + aux[j].x;
        if (EQ) {
            const f3 w = vbi - mk(av[j]);
            kb += G * dot3(xij, w);
            gx = gx + kk * Jmul(G, Gp, d, xij, w);
        } else {
            const float2 o = other_end<CUBIC, EQ>(G, Gp, d, xj4);
            const f3 vbj = mk(av[j]);
            kb += G * dot3(xij, vbi) - o.x * dot3(xij, vbj);
            gx = gx + kk * (Jmul(G, Gp, d, xij, vbi) - Jmul(o.x, o.y, d, xij, vbj));
        }
    }
    float kbb = 0.0f;
    f3 gb = f3z();
    const uint32_t cb = nb.cfb[i];
    for (uint32_t k = 0; k < cb; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.fb[slot];
        const f3 xib = xi - mk(bpos[j]);
        const float G = pG_fb[slot];
        const float vb = bvol[j];
        kbb += vb * G * dot3(xib, vbi);
        const float d = len3(xib);
        gb = gb + vb * Jmul(G, Gpk<CUBIC>(d, invH), d, xib, vbi);
    }
    const float hdt = -0.5f * P.dt;
    const float kbar = hdt * (P.V * kb + kbb);
    const float4 a4 = ax[i];
    const f3 dx = hdt * (P.V * gx + ai.x * gb);
    ax[i] = make_float4(a4.x + dx.x, a4.y + dx.y, a4.z + dx.z, a4.w);
    float db = 0.0f;
    if (ai.z != 0.0f) {
        const float idt2 = 1.0f / (P.dt * P.dt);
        const float ebar = kbar * alpha[i] * idt2;
        abar[i] += kbar * ai.y * idt2;
        db = ebar * P.dt;
        if (ai.w != 0.0f) rbar[i] += ebar / P.rho0;
    }
    dbar[i] = db;
}

// vbar of the velocity the iteration started from: += (d delta / d v)^T deltabar, plus the
// position terms of delta.
template <bool CUBIC, bool EQ>
__global__ void __launch_bounds__(kBlock) k_adj_dbar(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                   const float4* __restrict__ vel, const float4* __restrict__ bpos,
                                                   const float* __restrict__ bvol, const float4* __restrict__ bvel,
                                                   const float* __restrict__ pG_ff, const float* __restrict__ pG_fb,
                                                   const float* __restrict__ dbar, const uint32_t* __restrict__ L, uint32_t it,
                                                   float4* __restrict__ av, float4* __restrict__ ax) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    if (it >= L[sid]) return;
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    if (i >= P.end) return;
    const float4 xi4 = pos[i];
    const f3 xi = mk(xi4), vi = mk(vel[i]);
    const float invH = 1.0f / (xi4.w * 4.0f);
    const float di = dbar[i];
    f3 accv = f3z(), accx = f3z();
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.ff[slot];
        const float4 xj4 = pos[j];
        const f3 xij = xi - mk(xj4);
        const float G = pG_ff[slot];
        const float d = len3(xij);
        const float Gp = Gpk<CUBIC>(d, invH);
        const f3 vij = vi - mk(vel[j]);
        if (EQ) {
            const float s = P.V * (di + dbar[j]);
            accv = accv + (s * G) * xij;
            accx = accx + s * Jmul(G, Gp, d, xij, vij);
        } else {
            const float2 o = other_end<CUBIC, EQ>(G, Gp, d, xj4);
            const float si = P.V * di, sj = P.V * dbar[j];
            accv = accv + (si * G + sj * o.x) * xij;
            accx = accx + (si * Jmul(G, Gp, d, xij, vij) + sj * Jmul(o.x, o.y, d, xij, vij));
        }
    }
    const uint32_t cb = nb.cfb[i];
    for (uint32_t k = 0; k < cb; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.fb[slot];
        const f3 xib = xi - mk(bpos[j]);
        const float G = pG_fb[slot];
        const float s = bvol[j] * di;
        accv = accv + (s * G) * xib;
        const float d = len3(xib);
        accx = accx + s * Jmul(G, Gpk<CUBIC>(d, invH), d, xib, vi - mk(bvel[j]));
    }
    const float4 v4 = av[i], a4 = ax[i];
    av[i] = make_float4(v4.x + accv.x, v4.y + accv.y, v4.z + accv.z, v4.w);
    ax[i] = make_float4(a4.x + accx.x, a4.y + accx.y, a4.z + accx.z, a4.w);
}

// ----------------------------------------------------------------- implicit viscosity / ST solve
// rev[k*N + i] = position of i in the list of its k-th neighbour (the lists are symmetric);
// kEmpty (and *bad = 1) if it is missing.
__global__ void __launch_bounds__(kBlock) k_adj_rev_index(Scenes sc, Nbr nb, uint32_t* __restrict__ rev, uint32_t* bad) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = sc.N;
    if (i >= sc.P[sid].end) return;
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const uint32_t j = nb.ff[size_t(k) * N + i];
        const uint32_t cj = nb.cff[j];
        uint32_t m = kEmpty;
        for (uint32_t q = 0; q < cj; q++)
            if (nb.ff[size_t(q) * N + j] == i) {
                m = q;
                break;
            }
        if (m == kEmpty) atomicOr(bad, 1u);
        rev[size_t(k) * N + i] = m;
    }
}
// cabT[k*N + i] = the coefficients (a_ji, b_ji) of the pair seen from the neighbour j
__global__ void __launch_bounds__(kBlock) k_adj_transpose_coeffs(Scenes sc, Nbr nb, const uint32_t* __restrict__ rev,
                                                               const float2* __restrict__ cab, float2* __restrict__ cabT) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = sc.N;
    if (i >= sc.P[sid].end) return;
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t m = rev[slot];
        cabT[slot] = m == kEmpty ? make_float2(0.f, 0.f) : cab[size_t(m) * N + nb.ff[slot]];
    }
}

__device__ __forceinline__ f3 rec_pv(float4 a, float4 b) { return {a.w, b.x, b.y}; }
// Start of the transposed solve A_f^T lambda = b from lambda = 0: r = p = b, partial r.r.
__global__ void __launch_bounds__(kBlock) k_adj_cg_init(Scenes sc, const float4* __restrict__ pos, const float4* __restrict__ b,
                                                      float4* __restrict__ x, float4* __restrict__ r, float4* __restrict__ rec,
                                                      double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < sc.N) {
        const float4 p4 = pos[i];
        float4 b4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (i < sc.P[sid].end) {
            b4 = b[i];
            part = double(b4.x * b4.x) + double(b4.y * b4.y) + double(b4.z * b4.z);
        }
        x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        r[i] = make_float4(b4.x, b4.y, b4.z, 0.f);
        rec[2 * size_t(i)] = make_float4(p4.x, p4.y, p4.z, b4.x);
        rec[2 * size_t(i) + 1] = make_float4(b4.y, b4.z, 0.f, 0.f);
    }
    const double s = block_sum_d(part, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}
// Warm start: r = b - A_f^T x0 (x0 packed in recA), p = r into recB, x = x0; partials of r.r
// at partials[blk] and of b.b at partials[nblk + blk].
__global__ void __launch_bounds__(kBlock) k_adj_cg_init_warm(Scenes sc, Nbr nb, const float2* __restrict__ cab,
                                                           const float2* __restrict__ cabT, const float4* __restrict__ recA,
                                                           const float4* __restrict__ b, float4* __restrict__ x,
                                                           float4* __restrict__ r, float4* __restrict__ recB,
                                                           double* __restrict__ partials, uint32_t nblk) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    double prr = 0.0, pbb = 0.0;
    if (i < N) {
        const float4 ra = recA[2 * size_t(i)], rb = recA[2 * size_t(i) + 1];
        const f3 xi = mk(ra), x0 = rec_pv(ra, rb);
        f3 ri = f3z();
        if (i < P.end) {
            f3 acc = f3z();
            const uint32_t c = nb.cff[i];
            for (uint32_t k = 0; k < c; k++) {
                const size_t slot = size_t(k) * N + i;
                const uint32_t j = nb.ff[slot];
                const float2 ab = cab[slot], abT = cabT[slot];
                const float4 ja = recA[2 * size_t(j)], jb = recA[2 * size_t(j) + 1];
                const f3 xij = xi - mk(ja);
                const f3 pj = rec_pv(ja, jb);
                acc = acc + (ab.x * x0 - abT.x * pj) + (ab.y * dot3(x0, xij) - abT.y * dot3(pj, xij)) * xij;
            }
            const f3 Ax = x0 * (1.0f + P.cg_factor) + acc;
            const f3 bi = mk(b[i]);
            ri = bi - Ax;
            prr = double(ri.x * ri.x) + double(ri.y * ri.y) + double(ri.z * ri.z);
            pbb = double(bi.x * bi.x) + double(bi.y * bi.y) + double(bi.z * bi.z);
            x[i] = make_float4(x0.x, x0.y, x0.z, 0.0f);
        } else {
            x[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        r[i] = make_float4(ri.x, ri.y, ri.z, 0.0f);
        recB[2 * size_t(i)] = make_float4(xi.x, xi.y, xi.z, ri.x);
        recB[2 * size_t(i) + 1] = make_float4(ri.y, ri.z, 0.f, 0.f);
    }
    block_partial2(prr, pbb, partials, nblk);
}
// Start of the transposed loop: r.r, and the stop threshold -tol * b.b (relative to the right-
// hand side, so a good starting guess does not tighten it) or tol.
__global__ void __launch_bounds__(kBlock) k_adj_fold_start(Scenes sc, const double* __restrict__ partials, uint32_t nblk,
                                                         SolverDev* stv, uint32_t* active, double tol) {
    __shared__ double lds[kBlock / 64];
    const StepParams P = sc.P[blockIdx.x];
    SolverDev* st = stv + blockIdx.x;
    double a = 0.0, b = 0.0;
    for (uint32_t k = threadIdx.x; k < P.nblk; k += blockDim.x) {
        a += partials[P.blk0 + k];
        b += partials[nblk + P.blk0 + k];
    }
    a = block_sum_d(a, lds);
    b = block_sum_d(b, lds);
    if (threadIdx.x == 0) {
        st->rr = a;
        st->log[0] = a;
        st->l = 1u;
        st->thr = tol < 0.0 ? -tol * b : tol;
        if (a < st->thr || a == 0.0) {
            st->done = 1u;
            atomicSub(active, 1u);
        }
    }
}

// (A_f^T p)_i = (1+f) p_i + sum_j [ (a_ij p_i - a_ji p_j) + (b_ij (p_i.x_ij) - b_ji (p_j.x_ij)) x_ij ]
// and partials of p.Ap, for scenes whose solve is still running.
__global__ void __launch_bounds__(kBlock) k_adj_cg_applyT(Scenes sc, Nbr nb, const float2* __restrict__ cab,
                                                        const float2* __restrict__ cabT, const float4* __restrict__ rec,
                                                        float4* __restrict__ Ap, const SolverDev* st, double* __restrict__ partials) {
    __shared__ double lds[kBlock / 64];
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    if (st[sid].done) return;
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    double part = 0.0;
    if (i < P.end) {
        const float4 ra = rec[2 * size_t(i)], rb = rec[2 * size_t(i) + 1];
        const f3 xi = mk(ra), pi = rec_pv(ra, rb);
        f3 acc = f3z();
        const uint32_t c = nb.cff[i];
        for (uint32_t k = 0; k < c; k++) {
            const size_t slot = size_t(k) * N + i;
            const uint32_t j = nb.ff[slot];
            const float2 ab = cab[slot], abT = cabT[slot];
            const float4 ja = rec[2 * size_t(j)], jb = rec[2 * size_t(j) + 1];
            const f3 xij = xi - mk(ja);
            const f3 pj = rec_pv(ja, jb);
            acc = acc + (ab.x * pi - abT.x * pj) + (ab.y * dot3(pi, xij) - abT.y * dot3(pj, xij)) * xij;
        }
        const f3 o = pi * (1.0f + P.cg_factor) + acc;
        Ap[i] = make_float4(o.x, o.y, o.z, 0.0f);
        part = double(pi.x * o.x) + double(pi.y * o.y) + double(pi.z * o.z);
    }
    const double s = block_sum_d(part, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// The pair term of the residual seen from particle "i" (self) with neighbour "j" (other):
//   Q = a p + b (p.x) x + cs s x,   p = v*_i - v*_j,  u = v_in_i - v_in_j,  x = x_i - x_j,
// with a, b, s, g, w, c of the CG set-up (MPSSPHNonPressureUtil.cuh:467-648, restated in
// k_cg_setup), and its vector-Jacobian product with lambda_i: phi = lambda . Q.
struct PairVjp {
    f3 gx, gu;
    float rs, ro, fs, fo, mu, sigma;
};
template <bool CUBIC>
__device__ __forceinline__ PairVjp cg_pair_vjp(const StepParams& P, f3 x, f3 p, f3 u, float ri, float rj, float fi,
                                               float fj, float mi, float mj, float hi, f3 lam) {
    const float invH = 1.0f / hi;
    const float d = len3(x);
    const float W = Wk<CUBIC>(d, invH), G = Gk<CUBIC>(d, invH), STW = STWk(d, invH), STGs = STGk(d, invH);
    const f3 stG = STGs * x;
    const float mij = 0.5f * (mi + mj);
    const float invD = 2.0f / (ri + rj);
    const float K = (fi + fj) * invD * 0.5f * STW;
    const float inner = STW + P.dt * (dot3(u, stG) - K);
    const float g = inner * invD;
    const float s = mij * g;
    const float dd = d * d + 0.01f * hi * hi;
    const float c4 = 1.0f / (d + kFltEps);
    const float c1 = 0.5f * (mi + mj) / rj;
    const float c = c1 / dd * G * c4;
    const float cs = P.dt * P.sigma / mi;
    const float cx = P.xsph / P.dt;
    const float cv = 10.0f * P.rho0 * P.mu / ri;
    const float b = -P.dt * cv * c;  // a_ij itself is not needed by the reverse pass
    // reverse
    PairVjp o;
    const float Lx = dot3(lam, x), Lp = dot3(lam, p), px = dot3(p, x);
    const float abar = Lp, bbar = px * Lx;
    float sbar = cs * Lx;
    f3 gx = b * (Lx * p + px * lam) + (cs * s) * lam;
    sbar += P.dt * cs * abar;
    const float wbar = P.dt * cx * abar;
    o.sigma = P.dt * (P.dt / mi) * s * abar + (P.dt / mi) * s * Lx;
    const float cbar = -P.dt * cv * bbar;
    o.mu = -P.dt * (10.0f * P.rho0 / ri) * c * bbar;
    float rs = P.dt * cv * c * bbar / ri;
    const float gbar = mij * sbar;
    const float ibar = gbar * invD;
    float invDbar = gbar * inner;
    float STWbar = ibar;
    o.gu = (ibar * P.dt) * stG;
    const f3 stGbar = (ibar * P.dt) * u;
    const float Kbar = -ibar * P.dt;
    o.fs = Kbar * invD * 0.5f * STW;
    o.fo = o.fs;
    invDbar += Kbar * (fi + fj) * 0.5f * STW;
    STWbar += Kbar * (fi + fj) * invD * 0.5f;
    const float dinv = -invDbar * invD * invD * 0.5f;
    rs += dinv;
    float ro = dinv;
    const float STGsbar = dot3(stGbar, x);
    gx = gx + STGs * stGbar;
    const float Wbar = wbar * mj / rj;
    ro += -wbar * mj * W / (rj * rj);
    ro += -cbar * c / rj;
    float dbar = cbar * c * (-2.0f * d / dd) - cbar * c * c4;
    const float Gbar = cbar * c1 / dd * c4;
    dbar += Wbar * (WGk<CUBIC>(d, invH) * d) + Gbar * Gpk<CUBIC>(d, invH) + STWbar * (STGs * d) + STGsbar * STGp(d, invH);
    if (d > 0.0f) gx = gx + (dbar / d) * x;
    o.gx = gx;
    o.rs = rs;
    o.ro = ro;
    return o;
}

// thetabar = -lambda . dR/dtheta for the solve's inputs: vbar_in (out), positions (ax +=),
// densities (rbar +=), ST factors (fbar, out) and per-workgroup partials of mubar, sigmabar.
// Every pair is evaluated at both ends: from i (lambda_i) for i's own terms and mu/sigma, and
// from the neighbour (lambda_j) for the terms in which i is the neighbour.
template <bool CUBIC>
__global__ void __launch_bounds__(kBlock) k_adj_cg_grad(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                      const float4* __restrict__ vin, const float4* __restrict__ vstar,
                                                      const float* __restrict__ rho, const float* __restrict__ fst,
                                                      const float4* __restrict__ lam, float4* __restrict__ avin,
                                                      float4* __restrict__ ax, float* __restrict__ rbar,
                                                      float* __restrict__ fbar, double* __restrict__ partials, uint32_t nblk) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    double pmu = 0.0, psig = 0.0;
    if (i < P.end) {
        const float4 xi4 = pos[i], vi4 = vin[i];
        const f3 xi = mk(xi4), ui = mk(vi4), si = mk(vstar[i]), li = mk(lam[i]);
        const float hi = xi4.w * 4.0f, mi = vi4.w, ri = rho[i], fi = fst[i];
        f3 gx = f3z(), gu = f3z();
        float gr = 0.0f, gf = 0.0f;
        const uint32_t c = nb.cff[i];
        for (uint32_t k = 0; k < c; k++) {
            const uint32_t j = nb.ff[size_t(k) * N + i];
            const float4 xj4 = pos[j], vj4 = vin[j];
            const f3 xij = xi - mk(xj4), uij = ui - mk(vj4), pij = si - mk(vstar[j]);
            const float rj = rho[j], fj = fst[j], mj = vj4.w;
            const PairVjp A = cg_pair_vjp<CUBIC>(P, xij, pij, uij, ri, rj, fi, fj, mi, mj, hi, li);
            const PairVjp Bq =
                cg_pair_vjp<CUBIC>(P, neg(xij), neg(pij), neg(uij), rj, ri, fj, fi, mj, mi, xj4.w * 4.0f, mk(lam[j]));
            gx = gx - A.gx + Bq.gx;
            gu = gu - A.gu + Bq.gu;
            gr += -A.rs - Bq.ro;
            gf += -A.fs - Bq.fo;
            pmu -= double(A.mu);
            psig -= double(A.sigma);
        }
        const float f1 = 1.0f + P.cg_factor;
        avin[i] = make_float4(f1 * li.x + gu.x, f1 * li.y + gu.y, f1 * li.z + gu.z, 0.0f);
        const float4 a4 = ax[i];
        ax[i] = make_float4(a4.x + gx.x, a4.y + gx.y, a4.z + gx.z, a4.w);
        rbar[i] += gr;
        fbar[i] = gf;
    }
    block_partial2(pmu, psig, partials, nblk);
}
// The ST factor fst_i = sum_j m_j (v_in_i - v_in_j) . (G_ij x_ij): vbar_in and positions.
template <bool CUBIC, bool EQ>
__global__ void __launch_bounds__(kBlock) k_adj_fst(Scenes sc, Nbr nb, const float4* __restrict__ pos, const float4* __restrict__ vin,
                                                  const float* __restrict__ pG_ff, const float* __restrict__ fbar,
                                                  float4* __restrict__ avin, float4* __restrict__ ax) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    if (i >= P.end) return;
    const float4 xi4 = pos[i], vi4 = vin[i];
    const f3 xi = mk(xi4), ui = mk(vi4);
    const float invH = 1.0f / (xi4.w * 4.0f), mi = vi4.w, fi = fbar[i];
    f3 accv = f3z(), accx = f3z();
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.ff[slot];
        const float4 vj4 = vin[j], xj4 = pos[j];
        const f3 xij = xi - mk(xj4);
        const float G = pG_ff[slot];
        const float d = len3(xij);
        const float Gp = Gpk<CUBIC>(d, invH);
        const f3 uij = ui - mk(vj4);
        if (EQ) {
            const float s = fi * vj4.w + fbar[j] * mi;
            accv = accv + (s * G) * xij;
            accx = accx + s * Jmul(G, Gp, d, xij, uij);
        } else {
            const float2 o = other_end<CUBIC, EQ>(G, Gp, d, xj4);
            const float si = fi * vj4.w, sj = fbar[j] * mi;
            accv = accv + (si * G + sj * o.x) * xij;
            accx = accx + (si * Jmul(G, Gp, d, xij, uij) + sj * Jmul(o.x, o.y, d, xij, uij));
        }
    }
    const float4 v4 = avin[i], a4 = ax[i];
    avin[i] = make_float4(v4.x + accv.x, v4.y + accv.y, v4.z + accv.z, v4.w);
    ax[i] = make_float4(a4.x + accx.x, a4.y + accx.y, a4.z + accx.z, a4.w);
}
// acc[s] += sum of scene s's partials (two interleaved sets: [0, nblk) and [nblk, 2 nblk))
__global__ void __launch_bounds__(kBlock) k_adj_fold2(Scenes sc, const double* __restrict__ partials, uint32_t nblk,
                                                    double* __restrict__ acc0, double* __restrict__ acc1) {
    __shared__ double lds[kBlock / 64];
    const StepParams P = sc.P[blockIdx.x];
    double a = 0.0, b = 0.0;
    for (uint32_t k = threadIdx.x; k < P.nblk; k += blockDim.x) {
        a += partials[P.blk0 + k];
        b += partials[nblk + P.blk0 + k];
    }
    a = block_sum_d(a, lds);
    b = block_sum_d(b, lds);
    if (threadIdx.x == 0) {
        acc0[blockIdx.x] += a;
        acc1[blockIdx.x] += b;
    }
}

// ----------------------------------------------------------------- density and alpha
// g_i = sum_j V G_ij x_ij + sum_b V_b G_ib x_ib and abar = -alphabar / a^2 (alpha = 1/a).
__global__ void __launch_bounds__(kBlock) k_adj_gvec(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                   const float4* __restrict__ bpos, const float* __restrict__ bvol,
                                                   const float* __restrict__ pG_ff, const float* __restrict__ pG_fb,
                                                   const float* __restrict__ abar, float4* __restrict__ gv) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    if (i >= P.end) return;
    const f3 xi = mk(pos[i]);
    f3 gf = f3z(), gb = f3z();
    float s = 0.0f;
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const size_t slot = size_t(k) * N + i;
        const f3 grad = (P.V * pG_ff[slot]) * (xi - mk(pos[nb.ff[slot]]));
        gf = gf + grad;
        s += dot3(grad, grad);
    }
    const uint32_t cb = nb.cfb[i];
    for (uint32_t k = 0; k < cb; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.fb[slot];
        gb = gb + (bvol[j] * pG_fb[slot]) * (xi - mk(bpos[j]));
    }
    f3 g = f3z();
    g = g + gf;
    g = g + gb;
    float a = 0.0f;
    a += s;
    a += dot3(g, g);
    gv[i] = f4(g, a > kEps ? -abar[i] / (a * a) : 0.0f);
}
// positions from rhobar (rho_i = rho0 (V W0 + sum_j V W_ij + sum_b V_b W_ib)) and from abar
// (a_i = sum_j |V G_ij x_ij|^2 + |g_i|^2).
template <bool CUBIC, bool EQ>
__global__ void __launch_bounds__(kBlock) k_adj_density(Scenes sc, Nbr nb, const float4* __restrict__ pos,
                                                      const float4* __restrict__ bpos, const float* __restrict__ bvol,
                                                      const float* __restrict__ pG_ff, const float* __restrict__ pG_fb,
                                                      const float* __restrict__ rbar, const float4* __restrict__ gv,
                                                      float4* __restrict__ ax) {
    const uint32_t sid = sc.blk_scene[blockIdx.x];
    const StepParams P = sc.P[sid];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t N = P.N;
    if (i >= P.end) return;
    const float4 xi4 = pos[i];
    const f3 xi = mk(xi4);
    const float invH = 1.0f / (xi4.w * 4.0f);
    const float ri = rbar[i];
    const float4 g4 = gv[i];
    const f3 gi = mk(g4);
    const float ai = g4.w;
    const float V = P.V;
    f3 acc = f3z();
    const uint32_t c = nb.cff[i];
    for (uint32_t k = 0; k < c; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.ff[slot];
        const float4 xj4 = pos[j];
        const f3 xij = xi - mk(xj4);
        const float G = pG_ff[slot];
        const float d = len3(xij);
        const float Gp = Gpk<CUBIC>(d, invH);
        const float4 gj4 = gv[j];
        if (EQ) {
            const float WG = CUBIC ? G : WGk<CUBIC>(d, invH);
            acc = acc + (P.rho0 * V * (ri + rbar[j]) * WG) * xij;
            acc = acc + (2.0f * V * V * (ai + gj4.w) * G * (G + Gp * d)) * xij;
            acc = acc + (2.0f * V) * Jmul(G, Gp, d, xij, ai * gi - gj4.w * mk(gj4));
        } else {
            const float invHj = 1.0f / (xj4.w * 4.0f);
            const float2 o = other_end<CUBIC, EQ>(G, Gp, d, xj4);
            const float WGi = CUBIC ? G : WGk<CUBIC>(d, invH), WGj = CUBIC ? o.x : WGk<CUBIC>(d, invHj);
            acc = acc + (P.rho0 * V * (ri * WGi + rbar[j] * WGj)) * xij;
            acc = acc + (2.0f * V * V * (ai * G * (G + Gp * d) + gj4.w * o.x * (o.x + o.y * d))) * xij;
            acc = acc + (2.0f * V) * (Jmul(G, Gp, d, xij, ai * gi) - Jmul(o.x, o.y, d, xij, gj4.w * mk(gj4)));
        }
    }
    const uint32_t cb = nb.cfb[i];
    for (uint32_t k = 0; k < cb; k++) {
        const size_t slot = size_t(k) * N + i;
        const uint32_t j = nb.fb[slot];
        const f3 xib = xi - mk(bpos[j]);
        const float G = pG_fb[slot];
        const float d = len3(xib);
        const float vb = bvol[j];
        acc = acc + (P.rho0 * vb * ri * (CUBIC ? G : WGk<CUBIC>(d, invH))) * xib;
        acc = acc + (2.0f * ai * vb) * Jmul(G, Gpk<CUBIC>(d, invH), d, xib, gi);
    }
    const float4 a4 = ax[i];
    ax[i] = make_float4(a4.x + acc.x, a4.y + acc.y, a4.z + acc.z, a4.w);
}

// ----------------------------------------------------------------- slot order between steps
// The hash of a step moved slot src[s] of the previous step to slot s: send the adjoints back.
__global__ void __launch_bounds__(kBlock) k_adj_unpermute(uint32_t n, const uint32_t* __restrict__ src, const float4* __restrict__ ax,
                                                        const float4* __restrict__ av, const float4* __restrict__ al,
                                                        float4* __restrict__ ax_out, float4* __restrict__ av_out,
                                                        float4* __restrict__ al_out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t o = src[s];
    if (o == kEmpty) return;
    ax_out[o] = ax[s];
    av_out[o] = av[s];
    if (al) al_out[o] = al[s];
}

__global__ void __launch_bounds__(kBlock) k_adj_add_cotangent(uint32_t n, const uint32_t* __restrict__ perm,
                                                            const float* __restrict__ gx, const float* __restrict__ gv,
                                                            float4* __restrict__ adx, float4* __restrict__ adv) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t r = perm[s];
    if (r == kEmpty) return;
    const size_t o = 3 * size_t(r);
    float4 a = adx[s];
    a.x += gx[o];
    a.y += gx[o + 1];
    a.z += gx[o + 2];
    adx[s] = a;
    if (gv) {
        float4 b = adv[s];
        b.x += gv[o];
        b.y += gv[o + 1];
        b.z += gv[o + 2];
        adv[s] = b;
    }
}

}  // namespace dev

// ===================================================================== host launchers
namespace launch {
using namespace dev;
static inline unsigned nbk(size_t n) { return unsigned((n + kBlock - 1) / kBlock); }

void adj_advect(const Scenes& sc, const float4* pos, const float4* vin, const float4* ax, float4* av, hipStream_t s) {
    if (sc.N) launch_k(k_adj_advect, nbk(sc.N), kBlock, 0, s, sc, pos, vin, ax, av);
}
void adj_peval(int df, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel, const float4* bpos,
               const float* bvol, const float4* bvel, const float* rho, const float* alpha, const float* pG_ff,
               const float* pG_fb, float4* aux, hipStream_t s) {
    if (!sc.N) return;
    if (df) launch_k(k_adj_peval<true>, nbk(sc.N), kBlock, 0, s, sc, nb, pos, vel, bpos, bvol, bvel, rho, alpha, pG_ff, pG_fb, aux);
    else launch_k(k_adj_peval<false>, nbk(sc.N), kBlock, 0, s, sc, nb, pos, vel, bpos, bvol, bvel, rho, alpha, pG_ff, pG_fb, aux);
}
// f(cubic, eq) with both flags as compile-time constants: the kernel instance for the kernel set
// (cubic spline or poly6/spiky) and for equal radii or not
template <class F>
static void by_kind(AdjKind k, F&& f) {
    using T = std::true_type;
    using U = std::false_type;
    if (k.cubic) {
        if (k.eq) f(T{}, T{});
        else f(T{}, U{});
    } else {
        if (k.eq) f(U{}, T{});
        else f(U{}, U{});
    }
}

void adj_kbar(AdjKind k, int df, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* bpos, const float* bvol,
              const float* pG_ff, const float* pG_fb, const float4* aux, const float* alpha, const float4* av,
              const uint32_t* L, uint32_t it, float* dbar, float4* ax, float* abar, float* rbar, hipStream_t s) {
    if (!sc.N) return;
    by_kind(k, [&](auto C, auto E) {
        constexpr bool c = decltype(C)::value, e = decltype(E)::value;
        if (df)
            launch_k(k_adj_kbar<true, c, e>, nbk(sc.N), kBlock, 0, s, sc, nb, pos, bpos, bvol, pG_ff, pG_fb, aux, alpha, av, L, it,
                     dbar, ax, abar, rbar);
        else
            launch_k(k_adj_kbar<false, c, e>, nbk(sc.N), kBlock, 0, s, sc, nb, pos, bpos, bvol, pG_ff, pG_fb, aux, alpha, av, L,
                     it, dbar, ax, abar, rbar);
    });
}
void adj_dbar(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel, const float4* bpos,
              const float* bvol, const float4* bvel, const float* pG_ff, const float* pG_fb, const float* dbar,
              const uint32_t* L, uint32_t it, float4* av, float4* ax, hipStream_t s) {
    if (!sc.N) return;
    by_kind(k, [&](auto C, auto E) {
        launch_k(k_adj_dbar<decltype(C)::value, decltype(E)::value>, nbk(sc.N), kBlock, 0, s, sc, nb, pos, vel, bpos, bvol,
                 bvel, pG_ff, pG_fb, dbar, L, it, av, ax);
    });
}
void adj_rev_index(const Scenes& sc, const Nbr& nb, uint32_t* rev, uint32_t* bad, hipStream_t s) {
    if (sc.N) launch_k(k_adj_rev_index, nbk(sc.N), kBlock, 0, s, sc, nb, rev, bad);
}
void adj_transpose_coeffs(const Scenes& sc, const Nbr& nb, const uint32_t* rev, const float2* cab, float2* cabT, hipStream_t s) {
    if (sc.N) launch_k(k_adj_transpose_coeffs, nbk(sc.N), kBlock, 0, s, sc, nb, rev, cab, cabT);
}
void adj_cg_init(const Scenes& sc, const float4* pos, const float4* b, float4* x, float4* r, float4* rec, double* partials,
                 hipStream_t s) {
    if (sc.N) launch_k(k_adj_cg_init, nbk(sc.N), kBlock, 0, s, sc, pos, b, x, r, rec, partials);
}
void adj_cg_applyT(const Scenes& sc, const Nbr& nb, const float2* cab, const float2* cabT, const float4* rec, float4* Ap,
                   const SolverDev* st, double* partials, hipStream_t s) {
    if (sc.N) launch_k(k_adj_cg_applyT, nbk(sc.N), kBlock, 0, s, sc, nb, cab, cabT, rec, Ap, st, partials);
}
void adj_cg_grad(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vin, const float4* vstar,
                 const float* rho, const float* fst, const float4* lam, float4* avin, float4* ax, float* rbar, float* fbar,
                 double* partials, hipStream_t s) {
    if (!sc.N) return;
    const unsigned g = nbk(sc.N);
    if (k.cubic)
        launch_k(k_adj_cg_grad<true>, g, kBlock, 0, s, sc, nb, pos, vin, vstar, rho, fst, lam, avin, ax, rbar, fbar, partials, g);
    else
        launch_k(k_adj_cg_grad<false>, g, kBlock, 0, s, sc, nb, pos, vin, vstar, rho, fst, lam, avin, ax, rbar, fbar, partials, g);
}
void adj_fst(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vin, const float* pG_ff,
             const float* fbar, float4* avin, float4* ax, hipStream_t s) {
    if (!sc.N) return;
    by_kind(k, [&](auto C, auto E) {
        launch_k(k_adj_fst<decltype(C)::value, decltype(E)::value>, nbk(sc.N), kBlock, 0, s, sc, nb, pos, vin, pG_ff, fbar,
                 avin, ax);
    });
}
void adj_fold2(const Scenes& sc, const double* partials, double* acc0, double* acc1, hipStream_t s) {
    if (sc.N) launch_k(k_adj_fold2, sc.S, kBlock, 0, s, sc, partials, nbk(sc.N), acc0, acc1);
}
void adj_gvec(const Scenes& sc, const Nbr& nb, const float4* pos, const float4* bpos, const float* bvol, const float* pG_ff,
              const float* pG_fb, const float* abar, float4* gv, hipStream_t s) {
    if (sc.N) launch_k(k_adj_gvec, nbk(sc.N), kBlock, 0, s, sc, nb, pos, bpos, bvol, pG_ff, pG_fb, abar, gv);
}
void adj_density(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* bpos, const float* bvol,
                 const float* pG_ff, const float* pG_fb, const float* rbar, const float4* gv, float4* ax, hipStream_t s) {
    if (!sc.N) return;
    by_kind(k, [&](auto C, auto E) {
        launch_k(k_adj_density<decltype(C)::value, decltype(E)::value>, nbk(sc.N), kBlock, 0, s, sc, nb, pos, bpos, bvol,
                 pG_ff, pG_fb, rbar, gv, ax);
    });
}
void adj_unpermute(uint32_t n, const uint32_t* src, const float4* ax, const float4* av, const float4* al, float4* ax_out,
                   float4* av_out, float4* al_out, hipStream_t s) {
    if (n) launch_k(k_adj_unpermute, nbk(n), kBlock, 0, s, n, src, ax, av, al, ax_out, av_out, al_out);
}
void adj_cg_init_warm(const Scenes& sc, const Nbr& nb, const float2* cab, const float2* cabT, const float4* recA,
                      const float4* b, float4* x, float4* r, float4* recB, double* partials, hipStream_t s) {
    if (sc.N) launch_k(k_adj_cg_init_warm, nbk(sc.N), kBlock, 0, s, sc, nb, cab, cabT, recA, b, x, r, recB, partials, nbk(sc.N));
}
void adj_fold_start(const Scenes& sc, const double* partials, SolverDev* st, uint32_t* active, double tol, hipStream_t s) {
    if (sc.N) launch_k(k_adj_fold_start, sc.S, kBlock, 0, s, sc, partials, nbk(sc.N), st, active, tol);
}

void adj_add_cotangent(uint32_t n, const uint32_t* perm, const float* gx, const float* gv, float4* adx, float4* adv,
                       hipStream_t s) {
    if (n) launch_k(k_adj_add_cotangent, nbk(n), kBlock, 0, s, n, perm, gx, gv, adx, adv);
}

}  // namespace launch
}  // namespace mps
