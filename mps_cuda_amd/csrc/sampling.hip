// sampling.hip — triangle mesh -> Akinci boundary particles on gfx950 (setup path, FP64).
//
// Restates kernel::ParticleSampling (MPSParticleSamplingUtil.cu:14-133, .cuh:10-294):
// per face, owned vertices, owned edges (interior points at spacing l/(n+1) with
// n = ceil((l - 1.01 d)/d), d = R_m/2) and scan lines parallel to the shortest edge. One
// thread per face counts, the host scans the counts, one thread per face writes. Face
// ownership bits (which face samples each shared vertex and edge) are built here too.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

#pragma clang fp contract(off)

namespace mps {
namespace dev {

struct d3 {
    double x, y, z;
};
__device__ __forceinline__ d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dotd(d3 a, d3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ double lend(d3 a) { return sqrt(dotd(a, a)); }
__device__ __forceinline__ d3 crossd(d3 x, d3 y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
__device__ __forceinline__ d3 normalized(d3 v) { return v * (1.0 / sqrt(dotd(v, v))); }

// GetLineSamplingNum (MPSParticleSamplingUtil.cuh:12-18): ceil taken in float.
__device__ __forceinline__ uint32_t line_num(double l, double d) {
    const double x = (l - d * 1.01) / d;
    if (x <= 0.0) return 0u;
    return static_cast<uint32_t>(ceilf(static_cast<float>(x)));
}

struct Emit {
    double* pos;  // 3 per sample
    uint32_t* face;
    uint32_t cur, end;
    __device__ void put(d3 x, uint32_t f) {
        if (cur < end) {
            pos[3 * size_t(cur)] = x.x;
            pos[3 * size_t(cur) + 1] = x.y;
            pos[3 * size_t(cur) + 2] = x.z;
            face[cur] = f;
        }
        cur++;
    }
};

// GenerateLineParticle (:39-67); `d` is a float parameter that receives the new spacing.
__device__ void gen_line(d3 a, d3 b, float d, uint32_t f, Emit* e, uint32_t* count) {
    b = b - a;
    const double l = lend(b);
    const uint32_t num = line_num(l, double(d));
    if (count) *count += num;
    if (!e || num == 0u) return;
    d = static_cast<float>(l / static_cast<double>(num + 1u));
    b = b * (double(d) / l);
    for (uint32_t i = 0; i < num; i++) {
        a = a + b;
        e->put(a, f);
    }
}

// ComputeSamplingNum_kernel (:69-144) when e == nullptr, GenerateBoundaryParticle_kernel
// (:146-246) otherwise. The count pass uses the double spacing, the generate pass the
// float one, as in the reference.
__device__ void sample_face(const double* V, const uint32_t* F, const uint32_t* R, uint32_t id, double radius, Emit* e,
                            uint32_t* count, uint32_t* short_edge = nullptr) {
    const uint32_t f0 = F[3 * id], f1 = F[3 * id + 1], f2 = F[3 * id + 2];
    d3 ns[3] = {{V[3 * f0], V[3 * f0 + 1], V[3 * f0 + 2]}, {V[3 * f1], V[3 * f1 + 1], V[3 * f1 + 2]},
                {V[3 * f2], V[3 * f2 + 1], V[3 * f2 + 2]}};
    d3 es[3] = {ns[1] - ns[0], ns[2] - ns[1], ns[0] - ns[2]};
    const double ls[3] = {lend(es[0]), lend(es[1]), lend(es[2])};
    uint32_t iShort = 0u;
    if (ls[iShort] > ls[1]) iShort = 1u;
    if (ls[iShort] > ls[2]) iShort = 2u;
    if (short_edge) *short_edge = iShort;
    const double d = (radius + radius) * 0.25;
    const uint32_t rTri = R[id];
    uint32_t num = 0u;
    for (uint32_t i = 0; i < 3u; i++) {
        if ((rTri >> i) & 1u) {
            num++;
            if (e) e->put(ns[i], id);
        }
        if ((rTri >> (i + 3u)) & 1u) {
            if (e) gen_line(ns[i], ns[(i + 1u) % 3u], float(d), id, e, nullptr);
            else num += line_num(ls[i], d);
        }
    }
    const uint32_t ino0 = (iShort + 1u) % 3u, ino1 = (iShort + 2u) % 3u;
    ns[ino1] = ns[iShort];
    const d3 s = normalized(crossd(es[iShort], crossd(es[iShort], es[ino1])));
    es[ino1] = d3{-es[ino1].x, -es[ino1].y, -es[ino1].z};
    const double l = dotd(s, es[ino0]);
    const uint32_t nt = line_num(l, d);
    if (nt > 0u) {
        const double stride = (l / static_cast<double>(nt + 1u)) / l;
        for (uint32_t i = 0; i < nt; i++) {
            ns[ino0] = ns[ino0] + stride * es[ino0];
            ns[ino1] = ns[ino1] + stride * es[ino1];
            if (e) gen_line(ns[ino0], ns[ino1], float(d), id, e, nullptr);
            else num += line_num(lend(ns[ino0] - ns[ino1]), d);
        }
    }
    if (count) *count = num;
}

// Per face: sample count and shortest edge (kept by the mesh between samplings).
__global__ void k_sample_count(const double* V, const uint32_t* F, const uint32_t* R, uint32_t nt, double radius,
                               uint32_t* counts, uint32_t* shorts) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nt) return;
    uint32_t c = 0, sh = 0;
    sample_face(V, F, R, id, radius, nullptr, &c, &sh);
    counts[id] = c;
    if (shorts) shorts[id] = sh;
}

// Faces with gen[id] == 0 (gen may be null: every face) keep their previous particles and
// are skipped here.
__global__ void k_sample_generate(const double* V, const uint32_t* F, const uint32_t* R, uint32_t nt, double radius,
                                  const uint8_t* gen, const uint32_t* offs, double* pos, uint32_t* face, uint32_t* bad) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nt || (gen && !gen[id])) return;
    Emit e{pos, face, offs[id], offs[id + 1]};
    sample_face(V, F, R, id, radius, &e, nullptr);
    if (e.cur != e.end) atomicOr(bad, 1u);
}

// SetBarycentric_kernel (:248-294) and the float4 boundary record (x, y, z, r = R_m/4) for the
// slots written by k_sample_generate (face[id] < nt; other slots are left alone). The mass
// goes to bvel.w; bvel.xyz is zeroed, or kept when keep_vel (a re-sampling leaves the
// velocity stored at each index where it is).
__global__ void k_sample_finish(const double* V, const uint32_t* F, uint32_t nt, const double* vmass, uint32_t n,
                                const double* pos, const uint32_t* face, double radius, int keep_vel, float4* bpos,
                                float4* bvel, float2* bbcc, uint32_t* bface, float4* bcolor) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const uint32_t fi = face[id];
    if (fi >= nt) return;  // a kept particle, or a count/generate mismatch (reported by the host)
    const uint32_t f0 = F[3 * fi], f1 = F[3 * fi + 1], f2 = F[3 * fi + 2];
    const d3 n0{V[3 * f0], V[3 * f0 + 1], V[3 * f0 + 2]}, n1{V[3 * f1], V[3 * f1 + 1], V[3 * f1 + 2]},
        n2{V[3 * f2], V[3 * f2 + 1], V[3 * f2 + 2]};
    const d3 x{pos[3 * size_t(id)], pos[3 * size_t(id) + 1], pos[3 * size_t(id) + 2]};
    const d3 n20 = n0 - n2, n21 = n1 - n2;
    const double t0 = dotd(n20, n20), t1 = dotd(n21, n21), t2 = dotd(n20, n21);
    const double t3 = dotd(n20, x - n2), t4 = dotd(n21, x - n2);
    const double det = t0 * t1 - t2 * t2;
    double w0 = 0.0, w1 = 0.0;
    if (fabs(det) > 1.0e-20) {
        const double invDet = 1.0 / det;
        w0 = (+t1 * t3 - t2 * t4) * invDet;
        w1 = (-t2 * t3 + t0 * t4) * invDet;
    }
    w0 = fmin(fmax(w0, 0.0), 1.0);
    w1 = fmin(fmax(w1, 0.0), 1.0);
    const double m = vmass[f0] * w0 + vmass[f1] * w1 + vmass[f2] * (1.0 - w0 - w1);
    bpos[id] = make_float4(float(x.x), float(x.y), float(x.z), float(radius * 0.25));
    const float4 v = keep_vel ? bvel[id] : make_float4(0.f, 0.f, 0.f, 0.f);
    bvel[id] = make_float4(v.x, v.y, v.z, float(m));
    bbcc[id] = make_float2(float(w0), float(w1));
    bface[id] = fi;
    bcolor[id] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The particles of faces that were not generated anew move from their previous offset to the
// new one (GenerateBoundaryParticle_kernel's else branch, :230-245), one thread per new slot:
// position/radius, barycentric weights, face id and colour; mass and velocity too when
// prev_bvel is given (otherwise they stay with the index).
__global__ void k_sample_copy(uint32_t nt, const uint32_t* curr, const uint32_t* prev, const uint8_t* gen, uint32_t n,
                              const float4* prev_bpos, const float2* prev_bbcc, const uint32_t* prev_bface,
                              const float4* prev_bcolor, const float4* prev_bvel, float4* bpos, float2* bbcc,
                              uint32_t* bface, float4* bcolor, float4* bvel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t lo = 0, hi = nt;  // last face t with curr[t] <= i (its range holds i: curr[nt] = n > i)
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (curr[mid] <= i) lo = mid;
        else hi = mid;
    }
    if (gen[lo]) return;
    const uint32_t j = prev[lo] + (i - curr[lo]);
    bpos[i] = prev_bpos[j];
    bbcc[i] = prev_bbcc[j];
    bface[i] = prev_bface[j];
    bcolor[i] = prev_bcolor[j];
    if (prev_bvel) bvel[i] = prev_bvel[j];
}

// ---- face ownership (the result of RTriangleBuild_kernel, MPSMeshObject.cuh:41-101)
// Every vertex and every edge shared by several faces is sampled once, by its owner:
// - a face owns vertex v when no face containing v has a smaller id;
// - a face owns edge {a, b} when no face containing both a and b has a smaller id.
// Bits 0-2 of a face's word: owned vertices F[3t+i]; bits 3-5: owned edges
// (F[3t+i], F[3t+(i+1)%3]). A degenerate edge (a == b) is shared by every face containing a.
// The smallest face per vertex is an atomicMin; the smallest face per edge is an atomicMin in an
// open-addressed table keyed by the sorted vertex pair (load factor <= 1/2, so every probe ends).
constexpr unsigned long long kNoEdge = 0xFFFFFFFFFFFFFFFFull;
__device__ __forceinline__ uint32_t edge_slot(unsigned long long key, uint32_t mask) {
    return uint32_t((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}
__device__ __forceinline__ unsigned long long edge_key(uint32_t a, uint32_t b) {
    return a < b ? ((unsigned long long)a << 32) | b : ((unsigned long long)b << 32) | a;
}

__global__ void k_owner_min(const uint32_t* __restrict__ F, uint32_t nt, uint32_t* __restrict__ vmin,
                            unsigned long long* __restrict__ ekey, uint32_t* __restrict__ emin, uint32_t mask) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    for (uint32_t i = 0; i < 3u; i++) {
        const uint32_t a = F[3 * size_t(t) + i], b = F[3 * size_t(t) + (i == 2u ? 0u : i + 1u)];
        atomicMin(&vmin[a], t);
        if (a == b) continue;
        const unsigned long long key = edge_key(a, b);
        uint32_t sl = edge_slot(key, mask);
        for (uint32_t probe = 0; probe <= mask; probe++) {
            const unsigned long long prev = atomicCAS(&ekey[sl], kNoEdge, key);
            if (prev == kNoEdge || prev == key) {
                atomicMin(&emin[sl], t);
                break;
            }
            sl = (sl + 1u) & mask;
        }
    }
}

__global__ void k_owner_bits(const uint32_t* __restrict__ F, uint32_t nt, const uint32_t* __restrict__ vmin,
                             const unsigned long long* __restrict__ ekey, const uint32_t* __restrict__ emin, uint32_t mask,
                             uint32_t* __restrict__ R) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    uint32_t bits = 0u;
    for (uint32_t i = 0; i < 3u; i++) {
        const uint32_t a = F[3 * size_t(t) + i], b = F[3 * size_t(t) + (i == 2u ? 0u : i + 1u)];
        if (vmin[a] == t) bits |= 1u << i;
        uint32_t owner = vmin[a];  // a degenerate edge: every face containing a shares it
        if (a != b) {
            const unsigned long long key = edge_key(a, b);
            uint32_t sl = edge_slot(key, mask);
            for (uint32_t probe = 0; probe <= mask && ekey[sl] != key; probe++) sl = (sl + 1u) & mask;
            owner = emin[sl];
        }
        if (owner == t) bits |= 1u << (i + 3u);
    }
    R[t] = bits;
}

}  // namespace dev

namespace launch {
using namespace dev;
size_t owner_table_slots(uint32_t nt) {
    size_t slots = 1;
    while (slots < 6 * size_t(nt)) slots <<= 1;
    return slots;
}
void owner_build(const uint32_t* F, uint32_t nt, uint32_t nv, uint32_t* vmin, unsigned long long* ekey, uint32_t* emin,
                 size_t slots, uint32_t* R, hipStream_t s) {
    if (!nt) return;
    fill_u32(vmin, nv, 0xFFFFFFFFu, s);
    fill_u32(reinterpret_cast<uint32_t*>(ekey), 2 * slots, 0xFFFFFFFFu, s);
    fill_u32(emin, slots, 0xFFFFFFFFu, s);
    const uint32_t mask = uint32_t(slots - 1);
    launch_k(k_owner_min, (nt + 255) / 256, 256, 0, s, F, nt, vmin, ekey, emin, mask);
    launch_k(k_owner_bits, (nt + 255) / 256, 256, 0, s, F, nt, vmin, ekey, emin, mask, R);
}
void sample_count(const double* V, const uint32_t* F, const uint32_t* R, uint32_t nt, double radius, uint32_t* counts,
                  uint32_t* shorts, hipStream_t s) {
    if (nt) launch_k(k_sample_count, (nt + 63) / 64, 64, 0, s, V, F, R, nt, radius, counts, shorts);
}
void sample_generate(const double* V, const uint32_t* F, const uint32_t* R, uint32_t nt, double radius, const uint8_t* gen,
                     const uint32_t* offs, double* pos, uint32_t* face, uint32_t* bad, hipStream_t s) {
    if (nt) launch_k(k_sample_generate, (nt + 63) / 64, 64, 0, s, V, F, R, nt, radius, gen, offs, pos, face, bad);
}
void sample_finish(const double* V, const uint32_t* F, uint32_t nt, const double* vmass, uint32_t n, const double* pos,
                   const uint32_t* face, double radius, int keep_vel, float4* bpos, float4* bvel, float2* bbcc,
                   uint32_t* bface, float4* bcolor, hipStream_t s) {
    if (n)
        launch_k(k_sample_finish, (n + 255) / 256, 256, 0, s, V, F, nt, vmass, n, pos, face, radius, keep_vel, bpos, bvel,
                 bbcc, bface, bcolor);
}
void sample_copy(uint32_t nt, const uint32_t* curr, const uint32_t* prev, const uint8_t* gen, uint32_t n,
                 const float4* prev_bpos, const float2* prev_bbcc, const uint32_t* prev_bface, const float4* prev_bcolor,
                 const float4* prev_bvel, float4* bpos, float2* bbcc, uint32_t* bface, float4* bcolor, float4* bvel,
                 hipStream_t s) {
    if (n && nt)
        launch_k(k_sample_copy, (n + 255) / 256, 256, 0, s, nt, curr, prev, gen, n, prev_bpos, prev_bbcc, prev_bface,
                 prev_bcolor, prev_bvel, bpos, bbcc, bface, bcolor, bvel);
}
}  // namespace launch
}  // namespace mps
