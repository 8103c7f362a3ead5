// sph_oracle.cpp — CPU restatement of the MPS_CUDA SPH time step.
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker / CPU baseline; never by the product path.
//
// Every function cites the reference file:line it restates (paths relative to the
// reference's src/). Build: g++ -O2 -fopenmp -ffp-contract=off (no FMA contraction, no
// fast-math) so FP32 results follow the reference's expression order bit for bit.
#include "sph_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// ---------------------------------------------------------------- small vector type
template <class T>
struct V3 {
    T x, y, z;
};
template <class T>
inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
// glm::dot for vec3: tmp = a*b; (tmp.x + tmp.y) + tmp.z  (include/glm/detail/func_geometric.inl:46-53)
template <class T>
inline T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// glm::length = sqrt(dot(v, v))  (func_geometric.inl:8-14)
template <class T>
inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// glm::cross (func_geometric.inl:68-78)
template <class T>
inline V3<T> cross(V3<T> x, V3<T> y) {
    return {x.y * y.z - y.y * x.z, x.z * y.x - y.z * x.x, x.x * y.y - y.x * x.y};
}
// glm::normalize = v * inversesqrt(dot(v, v)), inversesqrt = 1/sqrt (func_geometric.inl:82-89)
template <class T>
inline V3<T> normalize(V3<T> v) { return v * (T(1) / std::sqrt(dot(v, v))); }

template <class T>
inline V3<T> ld3(const std::vector<T>& a, size_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
template <class T>
inline void st3(std::vector<T>& a, size_t i, V3<T> v) { a[3 * i] = v.x; a[3 * i + 1] = v.y; a[3 * i + 2] = v.z; }

// ---------------------------------------------------------------- smoothing kernels
// Constants exactly as written in MPSSPHKernel.cuh:9-15 (static_cast<REAL> of the literal).
template <class T>
struct KC {
    static constexpr T W = static_cast<T>(1.5666814710608447114749495456982);
    static constexpr T G = static_cast<T>(14.323944878270580219199538703526);
    static constexpr T CUBICW = static_cast<T>(2.5464790894703253723021402139602);
    static constexpr T CUBICG = static_cast<T>(15.278874536821952233812841283761);
    static constexpr T STW = static_cast<T>(7.2756545413437867780061148970292);
    static constexpr T STG = static_cast<T>(21.826963624031360334018344691088);
    static constexpr T EPS = static_cast<T>(1.0e-20);
};

// WKernel — MPSSPHKernel.cuh:28-45 (poly6 default, cubic behind USE_CUBIC_KERNEL)
template <class T>
inline T Wk(T dist, T invh, int cubic) {
    const T ratio = dist * invh;
    if (ratio >= T(1)) return T(0);
    if (cubic) {
        if (ratio >= T(0.5)) {
            const T temp = (T(1) - ratio) * invh;
            return KC<T>::CUBICW * T(2) * temp * temp * temp;
        }
        return KC<T>::CUBICW * invh * invh * invh * ((T(6) * ratio - T(6)) * ratio * ratio + T(1));
    }
    const T tmp = (T(1) - ratio * ratio) * invh;
    return KC<T>::W * tmp * tmp * tmp;
}
// GKernel — MPSSPHKernel.cuh:46-65; the gradient of W_ij is Gk * x_ij.
template <class T>
inline T Gk(T dist, T invh, int cubic) {
    const T ratio = dist * invh;
    if (ratio < KC<T>::EPS || ratio >= T(1)) return T(0);
    const T invh2 = invh * invh;
    if (cubic) {
        if (ratio > T(0.5)) {
            const T temp = (T(1) - ratio) * invh2;
            return -KC<T>::CUBICG / dist * temp * temp;
        }
        return KC<T>::CUBICG / dist * invh2 * invh2 * (T(3) * ratio - T(2)) * ratio;
    }
    const T tmp = (T(1) - ratio) * invh2;
    return -KC<T>::G / dist * tmp * tmp;
}
// The viscosity operator's pair factor G(d) / ((d^2 + h2) (d + FLT_EPSILON)) of
// ComputeCGViscosityAp_kernel (MPSSPHNonPressureUtil.cuh:617-618), regrouped so that it costs one
// division: the 1/d inside G (MPSSPHKernel.cuh:46-65) joins the two denominators,
//   Phi = (G d) / ((d (d^2 + h2)) (d + FLT_EPSILON)).
// It differs from the reference's left-to-right evaluation by FP32 rounding only; the GPU
// evaluates the same expression in the same order.
template <class T>
inline T Phik(T dist, T invh, T h2, int cubic) {
    const T ratio = dist * invh;
    if (ratio < KC<T>::EPS || ratio >= T(1)) return T(0);
    const T invh2 = invh * invh;
    T num;
    if (cubic) {
        if (ratio > T(0.5)) {
            const T temp = (T(1) - ratio) * invh2;
            num = -KC<T>::CUBICG * temp * temp;
        } else {
            num = KC<T>::CUBICG * invh2 * invh2 * (T(3) * ratio - T(2)) * ratio;
        }
    } else {
        const T tmp = (T(1) - ratio) * invh2;
        num = -KC<T>::G * tmp * tmp;
    }
    return num / ((dist * (dist * dist + h2)) * (dist + T(FLT_EPSILON)));
}
// STWKernel — MPSSPHKernel.cuh:67-75
template <class T>
inline T STWk(T dist, T invh) {
    const T ratio = std::max(dist * invh, T(0.5));
    if (ratio >= T(1)) return T(0);
    const T temp = (T(1) - ratio) * invh;
    return KC<T>::STW * temp * temp * temp;
}
// STGKernel — MPSSPHKernel.cuh:76-84
template <class T>
inline T STGk(T dist, T invh) {
    const T ratio = dist * invh;
    if (ratio <= T(0.5) || ratio >= T(1)) return T(0);
    const T temp = (T(1) - ratio) * invh * invh;
    return -KC<T>::STG / dist * temp * temp;
}

// ---------------------------------------------------------------- hashing
struct IV3 {
    int x, y, z;
};
// GetGridPos — MPSSpatialHash.cuh:9-19: trunc(x * (1.0/cell)) + H/2, the product in double.
inline IV3 grid_pos(double x, double y, double z, double cell, const uint32_t* H) {
    const double inv = 1.0 / cell;
    return {static_cast<int>(x * inv) + static_cast<int>(H[0] >> 1),
            static_cast<int>(y * inv) + static_cast<int>(H[1] >> 1),
            static_cast<int>(z * inv) + static_cast<int>(H[2] >> 1)};
}
// GetGridIndex — MPSSpatialHash.cuh:21-27 (wrap with &(H-1))
inline uint32_t grid_index(IV3 p, const uint32_t* H) {
    return ((static_cast<uint32_t>(p.z) & (H[2] - 1)) * H[1] + (static_cast<uint32_t>(p.y) & (H[1] - 1))) * H[0] +
           (static_cast<uint32_t>(p.x) & (H[0] - 1));
}
// GetGridZIndex — MPSSpatialHash.cuh:29-45 (adds H/2 again, 10-bit spread, clamp 1024)
inline uint32_t split_by3(uint32_t x) {
    if (x == 1024u) x--;
    x = (x | x << 16u) & 0b00000011000000000000000011111111u;
    x = (x | x << 8u) & 0b00000011000000001111000000001111u;
    x = (x | x << 4u) & 0b00000011000011000011000011000011u;
    x = (x | x << 2u) & 0b00001001001001001001001001001001u;
    return x;
}
inline uint32_t grid_zindex(IV3 p, const uint32_t* H) {
    const uint32_t x = static_cast<uint32_t>(p.x + static_cast<int>(H[0] >> 1u)) & (H[0] - 1u);
    const uint32_t y = static_cast<uint32_t>(p.y + static_cast<int>(H[1] >> 1u)) & (H[1] - 1u);
    const uint32_t z = static_cast<uint32_t>(p.z + static_cast<int>(H[2] >> 1u)) & (H[2] - 1u);
    return split_by3(x) | (split_by3(y) << 1u) | (split_by3(z) << 2u);
}

struct Hash {
    std::vector<uint32_t> key, id, start, end;
};
struct CSR {
    std::vector<uint32_t> idx, nei;
};

// UpdateHash — MPSSpatialHash.cu:53-77: fill start/end with 0xFFFFFFFF, key/id per particle
// (InitHash_kernel .cuh:83-98), sort_by_key (stable: ties keep ascending id), then
// ReorderHash_kernel (.cuh:115-145).
template <class T>
void update_hash(Hash& h, const std::vector<T>& pos, size_t n, double cell, const uint32_t* H) {
    const size_t C = size_t(H[0]) * H[1] * H[2];
    h.start.assign(C, kEmpty);
    h.end.assign(C, kEmpty);
    h.key.resize(n);
    h.id.resize(n);
    if (n == 0) return;
    std::vector<uint32_t> k(n);
    for (size_t i = 0; i < n; i++)
        k[i] = grid_index(grid_pos(double(pos[3 * i]), double(pos[3 * i + 1]), double(pos[3 * i + 2]), cell, H), H);
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    for (size_t s = 0; s < n; s++) {
        h.key[s] = k[perm[s]];
        h.id[s] = perm[s];
    }
    for (size_t s = 0; s < n; s++) {
        const uint32_t kk = h.key[s];
        if (s == 0 || h.key[s - 1] != kk) {
            h.start[kk] = uint32_t(s);
            if (s > 0) h.end[h.key[s - 1]] = uint32_t(s);
        }
        if (s == n - 1) h.end[kk] = uint32_t(n);
    }
}

// Research — MPSSpatialHash.cuh:47-80: 27 cells, z outer, then y, then x; ids in sorted order.
template <class Fn>
inline void research(const Hash& h, const uint32_t* H, double cell, double px, double py, double pz, Fn fn) {
    const IV3 p = grid_pos(px, py, pz, cell, H);
    IV3 q;
    for (q.z = p.z - 1; q.z <= p.z + 1; q.z++)
        for (q.y = p.y - 1; q.y <= p.y + 1; q.y++)
            for (q.x = p.x - 1; q.x <= p.x + 1; q.x++) {
                const uint32_t g = grid_index(q, H);
                uint32_t s = h.start[g];
                const uint32_t e = h.end[g];
                for (; s < e; s++) fn(h.id[s]);
            }
}

// BuildNeighorhood — MPSSpatialHash.cu:142-218 with ComputeNeighborhoodSize_kernel /
// BuildNeighborhood_kernel (.cuh:147-271): j is a neighbour of i when
// |x_i - x_j| < (4 r_i + 4 r_j) * 0.5; self excluded for same-set lists.
template <class T>
void build_csr(CSR& out, const std::vector<T>& pos, const std::vector<T>& rad, size_t n,
               const std::vector<T>& rpos, const std::vector<T>& rrad, const Hash& rh,
               const uint32_t* H, double cell, bool same) {
    std::vector<std::vector<uint32_t>> lists(n);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ii = 0; ii < (long long)n; ii++) {
        const size_t i = size_t(ii);
        const V3<T> xi = ld3(pos, i);
        const T ri = rad[i] * T(4);
        auto& L = lists[i];
        research(rh, H, cell, double(xi.x), double(xi.y), double(xi.z), [&](uint32_t j) {
            if (same && j == i) return;
            const V3<T> xj = ld3(rpos, j);
            const T rj = rrad[j] * T(4);
            const T dist = length(xi - xj);
            if (dist < (ri + rj) * T(0.5)) L.push_back(j);
        });
    }
    out.idx.assign(n + 1, 0u);
    for (size_t i = 0; i < n; i++) out.idx[i + 1] = out.idx[i] + uint32_t(lists[i].size());
    out.nei.resize(out.idx[n]);
    for (size_t i = 0; i < n; i++) std::copy(lists[i].begin(), lists[i].end(), out.nei.begin() + out.idx[i]);
}

// ---------------------------------------------------------------- mesh (setup, always FP64)
struct Mesh {
    std::vector<double> v;       // nv*3
    std::vector<uint32_t> f;     // nt*3
    std::vector<double> vmass;   // nv
    std::vector<uint32_t> rtri;  // nt
};

// MeshObject::buildAdjacency + RTriangleBuild_kernel (MPSMeshObject.cu:178-237,
// MPSMeshObject.cuh:41-101): per-vertex face lists sorted by (vertex, face); a face owns
// vertex i when it is the smallest face id at that vertex, and owns edge (i, i+1) when no
// other face sharing both vertices has a smaller id than the first such face found.
void rtri_build(size_t nv, const uint32_t* F, size_t nt, uint32_t* rtri) {
    std::vector<uint32_t> idx(nv + 1, 0u);
    for (size_t t = 0; t < nt; t++)
        for (int k = 0; k < 3; k++) idx[F[3 * t + k] + 1]++;
    for (size_t v = 0; v < nv; v++) idx[v + 1] += idx[v];
    std::vector<uint32_t> nb(idx[nv]);
    {
        std::vector<uint32_t> fill(idx.begin(), idx.end() - 1);
        for (size_t t = 0; t < nt; t++)
            for (int k = 0; k < 3; k++) nb[fill[F[3 * t + k]]++] = uint32_t(t);
        for (size_t v = 0; v < nv; v++) std::sort(nb.begin() + idx[v], nb.begin() + idx[v + 1]);
    }
    for (size_t id = 0; id < nt; id++) {
        const uint32_t* face = F + 3 * id;
        uint32_t info = 0u;
        uint32_t jStart = idx[face[0]], jEnd = idx[face[0] + 1];
        for (uint32_t i = 0, j = 1; i < 3u; i++, j = (j == 2u ? 0u : j + 1u)) {
            uint32_t ino = jStart;
            const uint32_t iEnd = jEnd;
            uint32_t iTri = nb[ino++];
            if (iTri == id) info |= (1u << i);
            jStart = idx[face[j]];
            jEnd = idx[face[j] + 1];
            uint32_t of = kEmpty;
            for (uint32_t jno = jStart; jno < jEnd; jno++) {
                if (nb[jno] == iTri && id != iTri) { of = iTri; break; }
            }
            for (; ino < iEnd && of == kEmpty; ino++) {
                iTri = nb[ino];
                for (uint32_t jno = jStart; jno < jEnd; jno++) {
                    if (nb[jno] == iTri && id != iTri) { of = iTri; break; }
                }
            }
            if (id < of) info |= (1u << (i + 3u));
        }
        rtri[id] = info;
    }
}

// GetLineSamplingNum — MPSParticleSamplingUtil.cuh:12-18 (ceil taken in float)
inline uint32_t line_num(double l, double d) {
    const double x = (l - d * 1.01) / d;
    if (x <= 0.0) return 0u;
    return static_cast<uint32_t>(std::ceil(static_cast<float>(x)));
}

struct Sample {
    V3<double> x;
    uint32_t face;
};

// GenerateLineParticle — MPSParticleSamplingUtil.cuh:39-67. `d` is a float parameter in
// the reference, and the recomputed spacing is stored back into that float.
inline void gen_line(V3<double> a, V3<double> b, float d, uint32_t face, std::vector<Sample>& out) {
    b = b - a;
    const double l = length(b);
    const uint32_t num = line_num(l, double(d));
    if (num == 0u) return;
    d = static_cast<float>(l / static_cast<double>(num + 1u));
    b = b * (double(d) / l);
    for (uint32_t i = 0; i < num; i++) {
        a = a + b;
        out.push_back({a, face});
    }
}

// Per-face sampling: ComputeSamplingNum_kernel (counts, .cuh:69-144) and
// GenerateBoundaryParticle_kernel (positions, .cuh:146-246).
void sample_face(const Mesh& M, size_t id, double radius, std::vector<Sample>* out, uint32_t* count,
                 uint32_t* short_edge = nullptr) {
    const uint32_t* face = &M.f[3 * id];
    V3<double> ns[3];
    for (int k = 0; k < 3; k++) ns[k] = {M.v[3 * face[k]], M.v[3 * face[k] + 1], M.v[3 * face[k] + 2]};
    V3<double> es[3] = {ns[1] - ns[0], ns[2] - ns[1], ns[0] - ns[2]};
    const double ls[3] = {length(es[0]), length(es[1]), length(es[2])};
    uint32_t iShort = 0u;
    if (ls[iShort] > ls[1]) iShort = 1u;
    if (ls[iShort] > ls[2]) iShort = 2u;
    if (short_edge) *short_edge = iShort;
    const double d = (radius + radius) * 0.25;
    const uint32_t rTri = M.rtri[id];
    uint32_t num = 0u;
    for (uint32_t i = 0; i < 3u; i++) {
        if ((rTri >> i) & 1u) {
            num++;
            if (out) out->push_back({ns[i], uint32_t(id)});
        }
        if ((rTri >> (i + 3u)) & 1u) {
            num += line_num(ls[i], d);
            if (out) gen_line(ns[i], ns[(i + 1u) % 3u], float(d), uint32_t(id), *out);
        }
    }
    const uint32_t ino0 = (iShort + 1u) % 3u, ino1 = (iShort + 2u) % 3u;
    ns[ino1] = ns[iShort];
    const V3<double> s = normalize(cross(es[iShort], cross(es[iShort], es[ino1])));
    es[ino1] = {-es[ino1].x, -es[ino1].y, -es[ino1].z};
    const double l = dot(s, es[ino0]);
    const uint32_t nt = line_num(l, d);
    if (nt > 0u) {
        const double stride = (l / static_cast<double>(nt + 1u)) / l;
        for (uint32_t i = 0; i < nt; i++) {
            ns[ino0] = ns[ino0] + stride * es[ino0];
            ns[ino1] = ns[ino1] + stride * es[ino1];
            num += line_num(length(ns[ino0] - ns[ino1]), d);
            if (out) gen_line(ns[ino0], ns[ino1], float(d), uint32_t(id), *out);
        }
    }
    if (count) *count = num;
}

// ---------------------------------------------------------------- the oracle context
struct Base {
    virtual ~Base() = default;
    std::string err;
    virtual int precision() const = 0;
    virtual int set_physics(const mps_physics* p) = 0;
    virtual int set_fluid_material(const mps_sph_material* m) = 0;
    virtual int upload_fluid(size_t n, const double* pos, const double* vel, const double* rad, const double* mass) = 0;
    virtual int upload_boundary(size_t n, const double* pos, const double* rad, const double* mass) = 0;
    virtual int sample_mesh(const double* verts, size_t nv, const uint32_t* tris, size_t nt, const double* center,
                            const double* size, const mps_mesh_material* m) = 0;
    virtual int resample_mesh(const double* verts, size_t nv, const mps_mesh_material* m, size_t* regenerated) = 0;
    virtual int boundary_objects(uint32_t k, const uint32_t* counts) = 0;
    virtual int step(mps_step_stats* st) = 0;
    virtual int run_phase(int32_t ph, mps_step_stats* st) = 0;
    virtual int array_info(int32_t id, size_t* count, size_t* eb) = 0;
    virtual int get_array(int32_t id, void* dst, size_t bytes) = 0;
    virtual int set_array(int32_t id, const void* src, size_t bytes) = 0;
    virtual int cg_operator(const void* p, void* out, double f) = 0;
    uint32_t frame = 0;
};

template <class T>
struct Ctx final : Base {
    mps_config cfg{};
    mps_physics phys{};
    // SPHMaterialParam as T; volume/mass derived in double (MPSSPHMaterial.cpp:13-19)
    T mR = 0, mRho0 = 1, mV = 0, mMass = 0, mMu = 0, mXsph = 0, mSigma = 0;
    int cubic = 0;
    size_t N = 0, B = 0;
    std::vector<T> x, v, r, m, rho, alpha, kappa, delta, fst, cgb, force, predv, prevv;
    std::vector<float> color;
    std::vector<T> bx, br, bvol, bm, bv, bbcc;
    std::vector<uint32_t> bface;
    std::vector<float> bcolor;
    // Other fluid objects of a coupled world, in ascending phase order, with this object's
    // ref neighbour list into each (SpatialHash::BuildNeighorhood(obj, ref, refHash),
    // MPSSpatialHash.cu:184-218): the "ref" overloads of the density, factor, delta and apply
    // kernels add one sum per ref object, after this object's own sum and before the walls'.
    struct RefPhase {
        const Ctx<T>* o;
        CSR list;
    };
    std::vector<RefPhase> refs;
    void build_ref_lists() {
        for (RefPhase& rf : refs) {
            build_csr(rf.list, x, r, N, rf.o->x, rf.o->r, rf.o->fh, H(), cell(), false);
            if (rf.o->N == 0) rf.list.idx.assign(N + 1, 0u), rf.list.nei.clear();
        }
    }
    // Wall objects: the walls may be several boundary objects (contiguous ranges, object ids in
    // bobj). Each object keeps its own Z-sort range; a wall's volume sums its own object first
    // and then every other object (the boundary "ref" overload, MPSSPHBaseUtil.cu:25-42,
    // .cuh:43-76), and the fluid sums over the walls one object after the other
    // (ComputeDensitySub(sph, nei, boundary) and the like, called once per object).
    static constexpr uint32_t kMaxObjects = 4;
    std::vector<uint32_t> bobj;
    uint32_t nobj = 1;
    int set_boundary_objects(uint32_t k, const uint32_t* counts) {
        if (k == 0 || k > kMaxObjects || !counts) { err = "1 to 4 wall objects"; return MPS_ERR_INVALID; }
        size_t total = 0;
        for (uint32_t o = 0; o < k; o++) total += counts[o];
        if (total != B) { err = "object sizes do not add up to the wall particles"; return MPS_ERR_INVALID; }
        size_t at = 0;
        for (uint32_t o = 0; o < k; o++)
            for (uint32_t q = 0; q < counts[o]; q++) bobj[at++] = o;
        nobj = k;
        mesh_ok = false;
        bb = CSR{};
        bb.idx.assign(B + 1, 0u);
        bh = Hash{};
        return MPS_OK;
    }
    // the mesh of the last sample_mesh and its per-face sampling state
    // (MeshObject::m_samplingParticleSize / m_shortestEdgeID, MPSMeshObject.cu:31-45)
    Mesh mesh;
    std::vector<uint32_t> mcount, mshort;
    bool mesh_ok = false;
    Hash fh, bh;
    CSR ff, fb, bb;

    int precision() const override { return sizeof(T) == 8 ? 64 : 32; }
    double cell() const { return cfg.cell_size > 0.0 ? double(T(cfg.cell_size)) : double(mR); }
    const uint32_t* H() const { return cfg.hash_size; }
    T dt() const { return T(phys.dt); }

    int set_physics(const mps_physics* p) override {
        phys = *p;
        return MPS_OK;
    }
    int set_fluid_material(const mps_sph_material* mm) override {
        const double R = mm->radius;
        const double V = R * R * R / 48.0 * 3.141592;
        mR = T(R);
        mRho0 = T(mm->density);
        mV = T(V);
        mMass = T(mm->density * V);
        mMu = T(mm->viscosity);
        mXsph = T(mm->xsph);
        mSigma = T(mm->surface_tension);
        return MPS_OK;
    }
    void resize_fluid(size_t n) {
        N = n;
        x.assign(3 * n, T(0)); v.assign(3 * n, T(0)); r.assign(n, T(0)); m.assign(n, T(0));
        rho.assign(n, T(0)); alpha.assign(n, T(0)); kappa.assign(n, T(0)); delta.assign(n, T(0));
        fst.assign(n, T(0)); cgb.assign(3 * n, T(0)); force.assign(3 * n, T(0));
        predv.assign(3 * n, T(0)); prevv.assign(3 * n, T(0));
        color.assign(4 * n, 0.f);
        ff = CSR{}; fb = CSR{};
        ff.idx.assign(n + 1, 0u); fb.idx.assign(n + 1, 0u);
        fh = Hash{};
    }
    int upload_fluid(size_t n, const double* pos, const double* vel, const double* rad, const double* mass) override {
        resize_fluid(n);
        for (size_t i = 0; i < 3 * n; i++) { x[i] = T(pos[i]); v[i] = vel ? T(vel[i]) : T(0); }
        for (size_t i = 0; i < n; i++) { r[i] = T(rad[i]); m[i] = mass ? T(mass[i]) : mMass; }
        // colour initialised as in the scene setup (MPSSystem.cpp:584)
        for (size_t i = 0; i < n; i++) { color[4 * i + 2] = 1.f; color[4 * i + 3] = 1.f; }
        frame = 0;
        return MPS_OK;
    }
    void resize_boundary(size_t n) {
        B = n;
        bx.assign(3 * n, T(0)); br.assign(n, T(0)); bvol.assign(n, T(0)); bm.assign(n, T(0));
        bv.assign(3 * n, T(0)); bbcc.assign(2 * n, T(0)); bface.assign(n, 0u); bcolor.assign(4 * n, 0.f);
        bobj.assign(n, 0u);
        nobj = 1;
        bb = CSR{}; bb.idx.assign(n + 1, 0u);
        bh = Hash{};
    }
    int upload_boundary(size_t n, const double* pos, const double* rad, const double* mass) override {
        mesh_ok = false;
        resize_boundary(n);
        for (size_t i = 0; i < 3 * n; i++) bx[i] = T(pos[i]);
        for (size_t i = 0; i < n; i++) { br[i] = T(rad[i]); bm[i] = mass ? T(mass[i]) : T(0); }
        return MPS_OK;
    }

    // MeshObject::LoadMesh (MPSMeshObject.cu:58-103) from in-memory arrays, then
    // kernel::ParticleSampling::ParticleSampling (MPSParticleSamplingUtil.cu:42-133).
    int sample_mesh(const double* verts, size_t nv, const uint32_t* tris, size_t nt, const double* center,
                    const double* size, const mps_mesh_material* mat) override {
        if (!verts || !tris || nv == 0 || nt == 0) { err = "empty mesh"; return MPS_ERR_INVALID; }
        for (size_t i = 0; i < 3 * nt; i++)
            if (tris[i] >= nv) { err = "triangle index out of range"; return MPS_ERR_INVALID; }
        Mesh M;
        M.v.assign(verts, verts + 3 * nv);
        M.f.assign(tris, tris + 3 * nt);
        if (center && size) {
            // AABB::AddPoint keeps its `else if` (MPSAABB.h:20-28); Translate (MPSMeshObject.cu:130-151)
            double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
            for (size_t i = 0; i < nv; i++)
                for (int k = 0; k < 3; k++) {
                    const double q = verts[3 * i + k];
                    if (mn[k] > q) mn[k] = q;
                    else if (mx[k] < q) mx[k] = q;
                }
            double oc[3], sc[3];
            for (int k = 0; k < 3; k++) {
                oc[k] = (mn[k] + mx[k]) * 0.5;
                const double os = mx[k] - mn[k];
                sc[k] = os > 1.0e-10 ? size[k] / os : 0.0;
            }
            for (size_t i = 0; i < nv; i++)
                for (int k = 0; k < 3; k++) M.v[3 * i + k] = (M.v[3 * i + k] - oc[k]) * sc[k] + center[k];
        }
        // InitFaceInfo: vertex mass = total area * density / nv (MPSMeshObject.cu:153-175)
        double area = 0.0;
        for (size_t t = 0; t < nt; t++) {
            const uint32_t* f = &M.f[3 * t];
            V3<double> a{M.v[3 * f[0]], M.v[3 * f[0] + 1], M.v[3 * f[0] + 2]};
            V3<double> b{M.v[3 * f[1]], M.v[3 * f[1] + 1], M.v[3 * f[1] + 2]};
            V3<double> c{M.v[3 * f[2]], M.v[3 * f[2] + 1], M.v[3 * f[2] + 2]};
            area += 0.5 * length(cross(b - a, c - a));
        }
        M.vmass.assign(nv, area * mat->density / double(nv));
        M.rtri.resize(nt);
        rtri_build(nv, M.f.data(), nt, M.rtri.data());
        std::vector<Sample> S;
        std::vector<uint32_t> counts(nt), shorts(nt);
        for (size_t t = 0; t < nt; t++) {
            const size_t before = S.size();
            sample_face(M, t, mat->radius, &S, &counts[t], &shorts[t]);
            if (S.size() - before != counts[t]) { err = "sample count mismatch between count and generate passes"; return MPS_ERR_STATE; }
        }
        resize_boundary(S.size());
        for (size_t i = 0; i < S.size(); i++) {
            bx[3 * i] = T(S[i].x.x); bx[3 * i + 1] = T(S[i].x.y); bx[3 * i + 2] = T(S[i].x.z);
            br[i] = T(mat->radius * 0.25);
            bface[i] = S[i].face;
            set_barycentric(M, i, S[i].x, S[i].face);
        }
        mesh = std::move(M);
        mcount = std::move(counts);
        mshort = std::move(shorts);
        mesh_ok = true;
        return MPS_OK;
    }

    // SetBarycentric_kernel (MPSParticleSamplingUtil.cuh:248-294), in double, for a particle
    // generated at x on face `face`: barycentric weights (clamped) and mass from the vertex masses.
    void set_barycentric(const Mesh& M, size_t i, V3<double> x, uint32_t face) {
        const uint32_t* f = &M.f[3 * face];
        V3<double> n0{M.v[3 * f[0]], M.v[3 * f[0] + 1], M.v[3 * f[0] + 2]};
        V3<double> n1{M.v[3 * f[1]], M.v[3 * f[1] + 1], M.v[3 * f[1] + 2]};
        V3<double> n2{M.v[3 * f[2]], M.v[3 * f[2] + 1], M.v[3 * f[2] + 2]};
        const V3<double> n20 = n0 - n2, n21 = n1 - n2;
        const double t0 = dot(n20, n20), t1 = dot(n21, n21), t2 = dot(n20, n21);
        const double t3 = dot(n20, x - n2), t4 = dot(n21, x - n2);
        const double det = t0 * t1 - t2 * t2;
        double w0 = 0.0, w1 = 0.0;
        if (std::fabs(det) > 1.0e-20) {
            const double invDet = 1.0 / det;
            w0 = (+t1 * t3 - t2 * t4) * invDet;
            w1 = (-t2 * t3 + t0 * t4) * invDet;
        }
        w0 = std::min(std::max(w0, 0.0), 1.0);
        w1 = std::min(std::max(w1, 0.0), 1.0);
        bbcc[2 * i] = T(w0);
        bbcc[2 * i + 1] = T(w1);
        bm[i] = T(M.vmass[f[0]] * w0 + M.vmass[f[1]] * w1 + M.vmass[f[2]] * (1.0 - w0 - w1));
    }

    // ParticleSampling called again on the same MeshObject after its vertices moved
    // (MPSParticleSamplingUtil.cu:14-133, .cuh:69-246). Per face, the count pass compares the
    // sample count and shortest edge with the ones kept from the previous call; a face where
    // either changed is generated anew (position, radius, face id, barycentric weights and
    // mass), every other face copies its previous particles, found by the exclusive scan of
    // the previous counts, to its new place. Nothing changes when no face changed. The
    // particle arrays are resized by index first (BoundaryParticleObject::Resize keeps the
    // leading elements, new ones are zero), and mass, velocity and volume are not moved with
    // the particles: a copied particle keeps the mass and velocity that were stored at its new
    // index. Without faithful quirks the copied particles carry their mass and velocity along.
    // The vertex masses of LoadMesh stay as they were.
    int boundary_objects(uint32_t k, const uint32_t* counts) override { return set_boundary_objects(k, counts); }
    int resample_mesh(const double* verts, size_t nv, const mps_mesh_material* mat, size_t* regenerated) override {
        if (!mesh_ok || nobj != 1) { err = "no mesh sampled on this context (mps_sample_boundary_mesh first)"; return MPS_ERR_STATE; }
        if (!verts || nv * 3 != mesh.v.size()) { err = "vertex count differs from the sampled mesh"; return MPS_ERR_INVALID; }
        const size_t nt = mesh.f.size() / 3;
        mesh.v.assign(verts, verts + 3 * nv);
        std::vector<uint32_t> prev_idx(nt + 1, 0u), curr_idx(nt + 1, 0u);
        std::vector<char> gen(nt, 0);
        size_t ngen = 0;
        for (size_t t = 0; t < nt; t++) {
            uint32_t c = 0, sh = 0;
            sample_face(mesh, t, mat->radius, nullptr, &c, &sh);
            prev_idx[t + 1] = prev_idx[t] + mcount[t];
            curr_idx[t + 1] = curr_idx[t] + c;
            if (sh != mshort[t] || c != mcount[t]) {
                gen[t] = 1;
                ngen++;
            }
            mshort[t] = sh;
            mcount[t] = c;
        }
        if (regenerated) *regenerated = ngen;
        if (ngen == 0) return MPS_OK;
        if (prev_idx[nt] != B) { err = "boundary particles no longer match the sampled mesh"; return MPS_ERR_STATE; }
        const std::vector<T> px = bx, pr = br, pbcc = bbcc, pm = bm, pv = bv;
        const std::vector<uint32_t> pface = bface;
        const std::vector<float> pcolor = bcolor;
        const size_t n = curr_idx[nt];
        B = n;
        bx.resize(3 * n, T(0)); br.resize(n, T(0)); bvol.resize(n, T(0)); bm.resize(n, T(0));
        bv.resize(3 * n, T(0)); bbcc.resize(2 * n, T(0)); bface.resize(n, 0u); bcolor.resize(4 * n, 0.f);
        bobj.assign(n, 0u);
        const bool faithful = cfg.faithful_quirks != 0;
        for (size_t t = 0; t < nt; t++) {
            size_t i = curr_idx[t];
            if (gen[t]) {
                std::vector<Sample> S;
                sample_face(mesh, t, mat->radius, &S, nullptr);
                if (S.size() != size_t(curr_idx[t + 1] - curr_idx[t])) { err = "sample count mismatch between count and generate passes"; return MPS_ERR_STATE; }
                for (const Sample& q : S) {
                    bx[3 * i] = T(q.x.x); bx[3 * i + 1] = T(q.x.y); bx[3 * i + 2] = T(q.x.z);
                    br[i] = T(mat->radius * 0.25);
                    bface[i] = q.face;
                    for (int k = 0; k < 4; k++) bcolor[4 * i + k] = 0.f;
                    set_barycentric(mesh, i, q.x, q.face);
                    i++;
                }
            } else {
                for (size_t j = prev_idx[t]; i < curr_idx[t + 1]; i++, j++) {
                    for (int k = 0; k < 3; k++) bx[3 * i + k] = px[3 * j + k];
                    br[i] = pr[j];
                    bface[i] = pface[j];
                    bbcc[2 * i] = pbcc[2 * j];
                    bbcc[2 * i + 1] = pbcc[2 * j + 1];
                    for (int k = 0; k < 4; k++) bcolor[4 * i + k] = pcolor[4 * j + k];
                    if (!faithful) {
                        bm[i] = pm[j];
                        for (int k = 0; k < 3; k++) bv[3 * i + k] = pv[3 * j + k];
                    }
                }
            }
        }
        bb = CSR{};
        bb.idx.assign(n + 1, 0u);
        bh = Hash{};
        return MPS_OK;
    }

    // ---------------------------------------------------------------- phases
    // SpatialHash::ZSort (MPSSpatialHash.cu:101-140): Morton key, stable sort of the payload.
    // Fluid: pos, mass, vel, colour, radius. Boundary adds faceID and bcc; the boundary
    // volume is NOT permuted (Q2).
    void zsort() {
        auto perm_of = [&](const std::vector<T>& pos, size_t n) {
            std::vector<uint32_t> k(n), p(n);
            for (size_t i = 0; i < n; i++)
                k[i] = grid_zindex(grid_pos(double(pos[3 * i]), double(pos[3 * i + 1]), double(pos[3 * i + 2]), cell(), H()), H());
            std::iota(p.begin(), p.end(), 0u);
            std::stable_sort(p.begin(), p.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
            return p;
        };
        auto apply = [](auto& arr, const std::vector<uint32_t>& p, size_t w) {
            auto old = arr;
            for (size_t s = 0; s < p.size(); s++)
                for (size_t k = 0; k < w; k++) arr[w * s + k] = old[w * p[s] + k];
        };
        if (B > 0) {
            auto p = perm_of(bx, B);
            if (nobj > 1)  // each object is Z-sorted within its own range (its own hash, MPSSystem.cpp:220-224)
                std::stable_sort(p.begin(), p.end(), [&](uint32_t a, uint32_t b) { return bobj[a] < bobj[b]; });
            apply(bx, p, 3); apply(bm, p, 1); apply(bv, p, 3); apply(bcolor, p, 4); apply(br, p, 1);
            apply(bface, p, 1); apply(bbcc, p, 2);
            if (!cfg.faithful_quirks) apply(bvol, p, 1);
        }
        if (N > 0) {
            const auto p = perm_of(x, N);
            apply(x, p, 3); apply(m, p, 1); apply(v, p, 3); apply(color, p, 4); apply(r, p, 1);
        }
    }
    void hash() {
        update_hash(bh, bx, B, cell(), H());
        update_hash(fh, x, N, cell(), H());
    }
    void neighbors() {
        build_csr(bb, bx, br, B, bx, br, bh, H(), cell(), true);
        build_csr(ff, x, r, N, x, r, fh, H(), cell(), true);
        build_csr(fb, x, r, N, bx, br, bh, H(), cell(), false);
        if (B == 0) fb.idx.assign(N + 1, 0u), fb.nei.clear();
    }
    template <class Fn>
    static void nsearch(const CSR& c, size_t i, Fn fn) {
        for (uint32_t k = c.idx[i]; k < c.idx[i + 1]; k++) fn(c.nei[k]);
    }
    // ComputeBoundaryParticleVolume_kernel + Final (MPSSPHBaseUtil.cuh:10-41,78-86):
    // V_b <- 1/(V_b + W(0) + sum W_bk)  (V_b is never cleared: Q1)
    void boundary_volume() {
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)B; ii++) {
            const size_t i = size_t(ii);
            const T hi = br[i] * T(4);
            const T invHi = T(1) / hi;
            const V3<T> xi = ld3(bx, i);
            const uint32_t own = bobj[i];
            T vo[kMaxObjects] = {};
            vo[own] = Wk(T(0), invHi, cubic);
            nsearch(bb, i, [&](uint32_t j) { vo[bobj[j]] += Wk(length(xi - ld3(bx, j)), invHi, cubic); });
            T vsum = cfg.faithful_quirks ? bvol[i] : T(0);
            vsum += vo[own];
            for (uint32_t o = 0; o < nobj; o++)
                if (o != own) vsum += vo[o];
            bvol[i] = T(1) / vsum;
        }
    }
    // ComputeDensity_kernel ff/fb + Final + DensityColorTest (MPSSPHBaseUtil.cuh:88-221)
    void density() {
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T hi = r[i] * T(4);
            const T invHi = T(1) / hi;
            const V3<T> xi = ld3(x, i);
            T sf = mV * Wk(T(0), invHi, cubic);
            nsearch(ff, i, [&](uint32_t j) { sf += mV * Wk(length(xi - ld3(x, j)), invHi, cubic); });
            T sb[kMaxObjects] = {};
            nsearch(fb, i, [&](uint32_t j) { sb[bobj[j]] += bvol[j] * Wk(length(xi - ld3(bx, j)), invHi, cubic); });
            T d = T(0);
            d += sf;
            for (const RefPhase& rf : refs) {  // ComputeDensity_kernel ref (MPSSPHBaseUtil.cuh:122-156)
                T sr = T(0);
                nsearch(rf.list, i, [&](uint32_t j) { sr += rf.o->mV * Wk(length(xi - ld3(rf.o->x, j)), invHi, cubic); });
                d += sr;
            }
            for (uint32_t o = 0; o < nobj; o++) d += sb[o];
            rho[i] = d * mRho0;
            const float ratio = static_cast<float>(rho[i] / mRho0);
            auto clampf = [](float a) { return std::min(std::max(a, 0.f), 1.f); };
            color[4 * i + 0] = clampf(ratio - 1.0f);
            color[4 * i + 1] = clampf(ratio < 1.0f ? ratio : 1.0f - (ratio - 1.0f));
            color[4 * i + 2] = clampf(1.0f - ratio);
            color[4 * i + 3] = 1.0f;
        }
    }
    // ComputeDFSPHFactor_kernel ff/fb + Final (MPSSPHDFSPHUtil.cuh:10-137):
    // g = sum V grad + sum V_b grad; s = sum |V grad|^2 (fluid only); alpha = 1/(s + g.g)
    void alpha_factor() {
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T hi = r[i] * T(4);
            const T invHi = T(1) / hi;
            const V3<T> xi = ld3(x, i);
            V3<T> gf{0, 0, 0};
            T s = T(0);
            nsearch(ff, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(x, j);
                const V3<T> grad = mV * Gk(length(xij), invHi, cubic) * xij;
                gf = gf + grad;
                s += dot(grad, grad);
            });
            V3<T> gb[kMaxObjects] = {};
            nsearch(fb, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(bx, j);
                gb[bobj[j]] = gb[bobj[j]] + bvol[j] * Gk(length(xij), invHi, cubic) * xij;
            });
            V3<T> g{0, 0, 0};
            g = g + gf;
            T a = T(0);
            a += s;
            for (const RefPhase& rf : refs) {  // ComputeDFSPHFactor_kernel ref (MPSSPHDFSPHUtil.cuh:49-88)
                V3<T> gr{0, 0, 0};
                T sr = T(0);
                nsearch(rf.list, i, [&](uint32_t j) {
                    const V3<T> xij = xi - ld3(rf.o->x, j);
                    const V3<T> grad = rf.o->mV * Gk(length(xij), invHi, cubic) * xij;
                    gr = gr + grad;
                    sr += dot(grad, grad);
                });
                g = g + gr;
                a += sr;
            }
            for (uint32_t o = 0; o < nobj; o++) g = g + gb[o];
            a += dot(g, g);
            alpha[i] = a > KC<T>::EPS ? T(1) / a : T(0);
        }
    }
    // ResetForce, ApplyGravity_kernel, UpdateVelocity_kernel (MPSAdvectUtil.cuh:8-34)
    void gravity() {
        const V3<T> g{T(phys.gravity[0]), T(phys.gravity[1]), T(phys.gravity[2])};
        for (size_t i = 0; i < N; i++) {
            const V3<T> F = V3<T>{0, 0, 0} + g * m[i];
            st3(force, i, F);
            if (m[i] > T(1.0e-10)) st3(v, i, ld3(v, i) + (dt() / m[i]) * F);
        }
    }
    // ComputeDensityDelta_kernel ff/fb (MPSSPHDFSPHUtil.cuh:139-257)
    void density_delta() {
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T hi = r[i] * T(4);
            const T invHi = T(1) / hi;
            const V3<T> xi = ld3(x, i), vi = ld3(v, i);
            T df = T(0);
            nsearch(ff, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(x, j);
                const V3<T> vij = vi - ld3(v, j);
                df += mV * Gk(length(xij), invHi, cubic) * dot(xij, vij);
            });
            T db[kMaxObjects] = {};
            nsearch(fb, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(bx, j);
                const V3<T> vij = vi - ld3(bv, j);
                db[bobj[j]] += bvol[j] * Gk(length(xij), invHi, cubic) * dot(xij, vij);
            });
            T d = T(0);
            d += df;
            for (const RefPhase& rf : refs) {  // ComputeDensityDelta_kernel ref (MPSSPHDFSPHUtil.cuh:177-216)
                T dr = T(0);
                nsearch(rf.list, i, [&](uint32_t j) {
                    const V3<T> xij = xi - ld3(rf.o->x, j);
                    const V3<T> vij = vi - ld3(rf.o->v, j);
                    dr += rf.o->mV * Gk(length(xij), invHi, cubic) * dot(xij, vij);
                });
                d += dr;
            }
            for (uint32_t o = 0; o < nobj; o++) d += db[o];
            delta[i] = d;
        }
    }
    // ComputeDFStiffness_kernel / ComputeCDStiffness_kernel (MPSSPHDFSPHUtil.cuh:259-340);
    // returns sum of positive errors in double.
    double stiffness(bool df) {
        double sum = 0.0;
        const T dtv = dt();
        for (size_t i = 0; i < N; i++) {
            T e;
            if (df) e = std::min(delta[i] * dtv, rho[i] / mRho0 + dtv * delta[i] - T(0.8));
            else e = rho[i] / mRho0 + dtv * delta[i] - T(1);
            if (e > T(0)) {
                sum += double(e);
                e *= alpha[i] / (dtv * dtv);
            } else {
                e = T(0);
            }
            kappa[i] = e;
        }
        return sum;
    }
    // ApplyDFSPH_kernel ff/fb + Final (MPSSPHDFSPHUtil.cuh:342-468), then UpdateVelocity
    void apply_pressure() {
        const T dtv = dt();
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T hi = r[i] * T(4);
            const T invHi = T(1) / hi;
            const V3<T> xi = ld3(x, i);
            const T pi = kappa[i];
            V3<T> ffo{0, 0, 0};
            nsearch(ff, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(x, j);
                const V3<T> fij = T(0.5) * mV * (pi + kappa[j]) * Gk(length(xij), invHi, cubic) * xij;
                ffo = ffo - fij;
            });
            V3<T> fbo[kMaxObjects] = {};
            nsearch(fb, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(bx, j);
                const V3<T> fij = T(0.5) * bvol[j] * pi * Gk(length(xij), invHi, cubic) * xij;
                fbo[bobj[j]] = fbo[bobj[j]] - fij;
            });
            V3<T> F{0, 0, 0};
            F = F + ffo;
            for (const RefPhase& rf : refs) {  // ApplyDFSPH_kernel ref (MPSSPHDFSPHUtil.cuh:379-418)
                V3<T> fro{0, 0, 0};
                nsearch(rf.list, i, [&](uint32_t j) {
                    const V3<T> xij = xi - ld3(rf.o->x, j);
                    const V3<T> fij = T(0.5) * rf.o->mV * (pi + rf.o->kappa[j]) * Gk(length(xij), invHi, cubic) * xij;
                    fro = fro - fij;
                });
                F = F + fro;
            }
            for (uint32_t o = 0; o < nobj; o++) F = F + fbo[o];
            F = m[i] * F;
            st3(force, i, F);
            if (m[i] > T(1.0e-10)) st3(v, i, ld3(v, i) + (dtv / m[i]) * F);
        }
    }
    // ComputeDFSPHDivergenceFree / ComputeDFSPHConstantDensity drivers
    // (MPSSPHDFSPHUtil.cu:270-441): l=1..cap; evaluate; break if E<tol && l>lmin; apply.
    uint32_t pressure_solve(bool df, double* log, uint32_t* hit_cap) {
        if (N == 0) { *hit_cap = 0; return 0; }
        const double tol = df ? cfg.df_tol : cfg.cd_tol;
        const uint32_t lmin = df ? cfg.df_min_iter : cfg.cd_min_iter;
        uint32_t l = 1u;
        *hit_cap = 1;
        while (l <= cfg.max_iter) {
            density_delta();
            const double E = stiffness(df) / double(N);
            if (log && l - 1 < MPS_MAX_ITER_LOG) log[l - 1] = E;
            if (E < tol && l > lmin) { *hit_cap = 0; break; }
            apply_pressure();
            l++;
        }
        return l;
    }
    // ComputeSurfaceTensionFactor_kernel (MPSSPHNonPressureUtil.cuh:425-465)
    void st_factor() {
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T invHi = T(1) / (r[i] * T(4));
            const V3<T> xi = ld3(x, i), vi = ld3(v, i);
            T d = T(0);
            nsearch(ff, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(x, j);
                const V3<T> vij = vi - ld3(v, j);
                const V3<T> grad = Gk(length(xij), invHi, cubic) * xij;
                d += m[j] * dot(vij, grad);
            });
            fst[i] = d;
        }
    }
    // g_ij of the surface-tension operator (MPSSPHNonPressureUtil.cuh:503-512, 563-572)
    inline T st_g(size_t i, uint32_t j, T invHi, V3<T> xij, T dist, T& mij) const {
        const V3<T> vij = ld3(v, i) - ld3(v, j);
        mij = T(0.5) * (m[i] + m[j]);
        const T invDij = T(2) / (rho[i] + rho[j]);
        const T stW = STWk(dist, invHi);
        const V3<T> stG = STGk(dist, invHi) * xij;
        return (stW + dt() * (dot(vij, stG) - (fst[i] + fst[j]) * invDij * T(0.5) * stW)) * invDij;
    }
    // Pair coefficients of the implicit operator, fixed for one solve (positions, v, rho and
    // Delta do not change while CG runs). With p_ij = p_i - p_j:
    //   (A p)_i = (1+f) p_i + sum_j [ a_ij p_ij + (b_ij (p_ij . x_ij)) x_ij ]
    //   a_ij = dt (cs_i s_ij + cx w_ij),  b_ij = -dt cv_i c_ij,
    //   s_ij = m_ij g_ij (surface tension), w_ij = (m_j/rho_j) W_ij (XSPH),
    //   c_ij = (0.5(m_i+m_j) (1/rho_j)) Phi(d_ij) (viscosity; the reference's
    //          (0.5(m_i+m_j)/rho_j) / (d^2 + 0.01 h_i^2) G_ij / (d + FLT_EPSILON), regrouped: Phik),
    //   cs_i = dt sigma/m_i, cx = XSPH/dt, cv_i = 10 rho0 mu/rho_i.
    // This is the operator of ComputeSurfaceTensionCGAp_kernel + ComputeCGViscosityAp_kernel
    // + ComputeFinalCGAp_kernel (MPSSPHNonPressureUtil.cuh:523-648) with the per-particle
    // scale factors distributed over the pair sums. The same pass forms the right-hand side
    // of ComputeSurfaceTensionCGb_kernel (:467-521): b_i = (dt sigma/m_i)(-sum s_ij x_ij) + v_i.
    std::vector<T> ca, cb, irho;
    void cg_setup() {
        const T dtv = dt();
        const T cx = mXsph / dtv;
        ca.assign(ff.nei.size(), T(0));
        cb.assign(ff.nei.size(), T(0));
        irho.assign(N, T(0));
        for (size_t i = 0; i < N; i++) irho[i] = T(1) / rho[i];
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T hi = r[i] * T(4);
            const T invHi = T(1) / hi;
            const T onePercentHSq = T(0.01) * hi * hi;
            const V3<T> xi = ld3(x, i);
            const T mi = m[i], di = rho[i];
            const T cs = dtv * mSigma / mi;
            const T cv = T(10) * mRho0 * mMu / di;
            V3<T> bi{0, 0, 0};
            for (uint32_t k = ff.idx[i]; k < ff.idx[i + 1]; k++) {
                const uint32_t j = ff.nei[k];
                const V3<T> xij = xi - ld3(x, j);
                const T dist = length(xij);
                T mij;
                const T gij = st_g(i, j, invHi, xij, dist, mij);
                const T sij = mij * gij;
                const T wij = m[j] / rho[j] * Wk(dist, invHi, cubic);
                const T cij = T(0.5) * (mi + m[j]) * irho[j] * Phik(dist, invHi, onePercentHSq, cubic);
                ca[k] = dtv * (cs * sij + cx * wij);
                cb[k] = -dtv * cv * cij;
                bi = bi - sij * xij;
            }
            st3(cgb, i, dt() * mSigma / m[i] * bi + ld3(v, i));
        }
    }
    void cg_apply(const std::vector<T>& p, std::vector<T>& Ap, T f) {
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const V3<T> xi = ld3(x, i), pi = ld3(p, i);
            V3<T> acc{0, 0, 0};
            for (uint32_t k = ff.idx[i]; k < ff.idx[i + 1]; k++) {
                const uint32_t j = ff.nei[k];
                const V3<T> xij = xi - ld3(x, j);
                const V3<T> pij = pi - ld3(p, j);
                acc = acc + (ca[k] * pij + (cb[k] * dot(pij, xij)) * xij);
            }
            st3(Ap, i, pi * (T(1) + f) + acc);
        }
    }
    static double dotd(const std::vector<T>& a, const std::vector<T>& b) {
        double s = 0.0;
        for (size_t k = 0; k < a.size(); k++) s += double(T(a[k] * b[k]));
        return s;
    }
    // ApplyImplicitViscosityNSurfaceTension + CGSolver (MPSSPHNonPressureUtil.cu:127-188,
    // MPSSolver.cuh:12-73, MPSCGKernel.cuh:5-133)
    uint32_t viscosity_cg(double* log, uint32_t* hit_cap) {
        *hit_cap = 0;
        if (N == 0) return 0;
        predv = v;
        st_factor();
        cg_setup();
        const size_t n3 = 3 * N;
        std::vector<T> rr(n3), p(n3), Ap(n3);
        cg_apply(predv, Ap, T(0));
        for (size_t k = 0; k < n3; k++) { rr[k] = cgb[k] - Ap[k]; p[k] = rr[k]; }
        uint32_t l = 1u;
        *hit_cap = 1;
        // a negative tolerance is relative to the initial r.r (a zero residual then ends the loop)
        const double thr = cfg.cg_tol < 0.0 ? -cfg.cg_tol * dotd(rr, rr) : cfg.cg_tol;
        while (l <= cfg.cg_max_iter) {
            cg_apply(p, Ap, T(cfg.cg_factor));
            const double r2 = dotd(rr, rr);
            const double pAp = dotd(p, Ap);
            if (log && l - 1 < MPS_MAX_ITER_LOG) log[l - 1] = r2;
            if (r2 < thr || (cfg.cg_tol < 0.0 && r2 == 0.0)) { *hit_cap = 0; break; }
            const T a = T(r2 / pAp);
            for (size_t k = 0; k < n3; k++) { predv[k] = predv[k] + a * p[k]; rr[k] = rr[k] - a * Ap[k]; }
            const double r2n = dotd(rr, rr);
            const T b = T(r2n / r2);
            for (size_t k = 0; k < n3; k++) p[k] = rr[k] + b * p[k];
            l++;
        }
        v = predv;
        return l;
    }

    // ---- Jacobi implicit viscosity (dead reference code, SPHUtil.cu:344-453 and the kernels
    // at MPSSPHNonPressureUtil.cuh:90-139,140-194,195-333,413-423); Volume -> material V,
    // material.radius -> h_i = 4 r_i, Density -> rho_i.
    // Per particle: A_i = I - a_i sum A_ij, rhs_i = v0_i - a_i sum A_ij u_j with
    // A_ij = V G/((d^2+0.01h^2)(d+eps)) x_ij x_ij^T and a_i = dt*10*mu/rho_i; u_new = rhs A^-1.
    void jacobi_candidate(const std::vector<T>& u, std::vector<T>& out) {
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T h = r[i] * T(4);
            const T invH = T(1) / h;
            const T onePercentHSq = T(0.01) * h * h;
            const V3<T> xi = ld3(x, i);
            T A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  // A[col][row] like glm
            V3<T> Avj{0, 0, 0};
            nsearch(ff, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(x, j);
                const T dist = length(xij);
                if (dist < h) {
                    const V3<T> uj = ld3(u, j);
                    const T c = mV * Gk(dist, invH, cubic) / ((dist * dist + onePercentHSq) * (dist + T(FLT_EPSILON)));
                    const T M[3][3] = {{c * (xij.x * xij.x), c * (xij.y * xij.x), c * (xij.z * xij.x)},
                                       {c * (xij.x * xij.y), c * (xij.y * xij.y), c * (xij.z * xij.y)},
                                       {c * (xij.x * xij.z), c * (xij.y * xij.z), c * (xij.z * xij.z)}};
                    // row vector * matrix (glm): component k = dot(u, column k)
                    const V3<T> uM{dot(uj, V3<T>{M[0][0], M[0][1], M[0][2]}),
                                   dot(uj, V3<T>{M[1][0], M[1][1], M[1][2]}),
                                   dot(uj, V3<T>{M[2][0], M[2][1], M[2][2]})};
                    Avj = Avj - uM;
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) A[a][b] = A[a][b] - M[a][b];
                }
            });
            const T al = dt() * T(10) * mMu / rho[i];
            Avj = Avj * al;
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) A[a][b] = A[a][b] * al + (a == b ? T(1) : T(0));
            // glm::inverse of a 3x3 (func_matrix.inl compute_inverse<3,3>)
            const T oneOverDet = T(1) / (+A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]) -
                                         A[1][0] * (A[0][1] * A[2][2] - A[2][1] * A[0][2]) +
                                         A[2][0] * (A[0][1] * A[1][2] - A[1][1] * A[0][2]));
            T I[3][3];
            I[0][0] = +(A[1][1] * A[2][2] - A[2][1] * A[1][2]) * oneOverDet;
            I[1][0] = -(A[1][0] * A[2][2] - A[2][0] * A[1][2]) * oneOverDet;
            I[2][0] = +(A[1][0] * A[2][1] - A[2][0] * A[1][1]) * oneOverDet;
            I[0][1] = -(A[0][1] * A[2][2] - A[2][1] * A[0][2]) * oneOverDet;
            I[1][1] = +(A[0][0] * A[2][2] - A[2][0] * A[0][2]) * oneOverDet;
            I[2][1] = -(A[0][0] * A[2][1] - A[2][0] * A[0][1]) * oneOverDet;
            I[0][2] = +(A[0][1] * A[1][2] - A[1][1] * A[0][2]) * oneOverDet;
            I[1][2] = -(A[0][0] * A[1][2] - A[1][0] * A[0][2]) * oneOverDet;
            I[2][2] = +(A[0][0] * A[1][1] - A[1][0] * A[0][1]) * oneOverDet;
            if (std::isinf(I[0][0]) || std::isnan(I[0][0]))
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) I[a][b] = T(0);
            const V3<T> rhs = ld3(v, i) + Avj;
            st3(out, i, V3<T>{dot(rhs, V3<T>{I[0][0], I[0][1], I[0][2]}), dot(rhs, V3<T>{I[1][0], I[1][1], I[1][2]}),
                              dot(rhs, V3<T>{I[2][0], I[2][1], I[2][2]})});
        }
    }
    // sum_i |u_i - a_i sum_j (x_ij.(u_i-u_j)) grad_ij - v0_i|  (ComputeJacobiError kernels)
    double jacobi_error(const std::vector<T>& u) {
        std::vector<T> e(N);
#pragma omp parallel for
        for (long long ii = 0; ii < (long long)N; ii++) {
            const size_t i = size_t(ii);
            const T h = r[i] * T(4);
            const T invH = T(1) / h;
            const T onePercentHSq = T(0.01) * h * h;
            const V3<T> xi = ld3(x, i), ui = ld3(u, i);
            V3<T> Av{0, 0, 0};
            nsearch(ff, i, [&](uint32_t j) {
                const V3<T> xij = xi - ld3(x, j);
                const T dist = length(xij);
                if (dist < h) {
                    const V3<T> grad = mV * Gk(dist, invH, cubic) / ((dist * dist + onePercentHSq) * (dist + T(FLT_EPSILON))) * xij;
                    Av = Av + dot(xij, ui - ld3(u, j)) * grad;
                }
            });
            Av = Av * (dt() * T(10) * mMu / rho[i]);
            e[i] = length(ui - Av - ld3(v, i));
        }
        double s = 0.0;
        for (size_t i = 0; i < N; i++) s += double(e[i]);
        return s;
    }
    uint32_t viscosity_jacobi(double* log, uint32_t* hit_cap) {
        *hit_cap = 0;
        if (N == 0) return 0;
        predv = v;
        prevv = v;
        std::vector<T> tmp[3] = {std::vector<T>(3 * N), std::vector<T>(3 * N), std::vector<T>(3 * N)};
        const double underRelax = 0.9, delta = 0.005;
        double omega = 1.0, rhoS = 0.995;
        double om[3] = {1.0, 1.0, 1.0};
        uint32_t l = 0u;
        *hit_cap = 1;
        while (l < cfg.jacobi_max_iter) {
            uint32_t minID = 0u;
            double errs[3] = {0.0, 0.0, 0.0};
            if (l < 10u) {
                jacobi_candidate(predv, tmp[0]);
                errs[0] = jacobi_error(tmp[0]);
            } else {
                if (l == 10u) {
                    om[0] = 2.0 / (2.0 - rhoS * rhoS);
                    om[1] = 2.0 / (2.0 - (rhoS + delta) * (rhoS + delta));
                    om[2] = 2.0 / (2.0 - (rhoS - delta) * (rhoS - delta));
                } else {
                    om[0] = 4.0 / (4.0 - rhoS * rhoS * omega);
                    om[1] = 4.0 / (4.0 - (rhoS + delta) * (rhoS + delta) * omega);
                    om[2] = 4.0 / (4.0 - (rhoS - delta) * (rhoS - delta) * omega);
                }
                std::vector<T> cand(3 * N);
                jacobi_candidate(predv, cand);
                const T ur = T(underRelax);
                for (int k = 0; k < 3; k++) {
                    const T w = T(om[k]);
                    for (size_t q = 0; q < 3 * N; q++)
                        tmp[k][q] = w * (ur * (cand[q] - predv[q]) + predv[q] - prevv[q]) + prevv[q];
                    errs[k] = jacobi_error(tmp[k]);
                }
                if (errs[minID] > errs[1]) minID = 1u;
                if (errs[minID] > errs[2]) minID = 2u;
                if (minID == 1u) rhoS += delta;
                else if (minID == 2u) rhoS -= delta;
                rhoS = std::min(std::max(0.5 + delta, rhoS), 1.0 - delta);
                omega = double(T(om[minID]));
            }
            prevv = predv;
            predv = tmp[minID];
            const double e = errs[minID] * (double(dt()) / double(N));
            if (log && l < MPS_MAX_ITER_LOG) log[l] = e;
            if (e < cfg.jacobi_tol) { *hit_cap = 0; break; }
            l++;
        }
        v = predv;
        return l;
    }
    // BoundaryCollision_kernel + UpdatePosition_kernel (MPSAdvectUtil.cuh:36-77)
    void advect() {
        const T dtv = dt();
        const T mn[3] = {T(phys.box_min[0]), T(phys.box_min[1]), T(phys.box_min[2])};
        const T mx[3] = {T(phys.box_max[0]), T(phys.box_max[1]), T(phys.box_max[2])};
        for (size_t i = 0; i < N; i++) {
            T vel[3] = {v[3 * i], v[3 * i + 1], v[3 * i + 2]};
            const T pos[3] = {x[3 * i] + vel[0] * dtv, x[3 * i + 1] + vel[1] * dtv, x[3 * i + 2] + vel[2] * dtv};
            for (int k = 0; k < 3; k++) {
                if (pos[k] < mn[k] && vel[k] < T(0)) vel[k] = (std::fabs(vel[k]) < T(1.0e-4)) ? T(0) : -vel[k] * T(0.4);
                else if (pos[k] > mx[k] && vel[k] > T(0)) vel[k] = (std::fabs(vel[k]) < T(1.0e-4)) ? T(0) : -vel[k] * T(0.4);
            }
            for (int k = 0; k < 3; k++) {
                v[3 * i + k] = vel[k];
                x[3 * i + k] = x[3 * i + k] + dtv * vel[k];
            }
        }
    }
    int cg_operator(const void* pin, void* pout, double f) override {
        st_factor();
        cg_setup();
        std::vector<T> p(static_cast<const T*>(pin), static_cast<const T*>(pin) + 3 * N), Ap(3 * N);
        cg_apply(p, Ap, T(f));
        std::memcpy(pout, Ap.data(), sizeof(T) * 3 * N);
        return MPS_OK;
    }
    bool finite_state() const {
        for (size_t k = 0; k < 3 * N; k++)
            if (!std::isfinite(double(x[k])) || !std::isfinite(double(v[k]))) return false;
        return true;
    }

    int run_phase(int32_t ph, mps_step_stats* st) override {
        mps_step_stats dummy;
        if (!st) st = &dummy;
        switch (ph) {
            case MPS_PHASE_ZSORT: zsort(); break;
            case MPS_PHASE_HASH: hash(); break;
            case MPS_PHASE_NEIGHBORS: neighbors(); break;
            case MPS_PHASE_BOUNDARY_VOLUME: boundary_volume(); break;
            case MPS_PHASE_DENSITY: density(); break;
            case MPS_PHASE_ALPHA: alpha_factor(); break;
            case MPS_PHASE_GRAVITY: gravity(); break;
            case MPS_PHASE_DIVERGENCE: st->l_df = pressure_solve(true, st->err_df, &st->hit_cap_df); break;
            case MPS_PHASE_VISCOSITY:
                if (cfg.visc_solver == MPS_VISC_CG) st->l_cg = viscosity_cg(st->rr_cg, &st->hit_cap_cg);
                else if (cfg.visc_solver == MPS_VISC_JACOBI) st->l_cg = viscosity_jacobi(st->rr_cg, &st->hit_cap_cg);
                break;
            case MPS_PHASE_CONSTANT_DENSITY: st->l_cd = pressure_solve(false, st->err_cd, &st->hit_cap_cd); break;
            case MPS_PHASE_ADVECT: advect(); break;
            default: err = "unknown phase"; return MPS_ERR_INVALID;
        }
        return MPS_OK;
    }
    // System::OnUpdate order (MPSSystem.cpp:191-291)
    int step(mps_step_stats* st) override {
        mps_step_stats dummy;
        if (!st) st = &dummy;
        std::memset(st, 0, sizeof(*st));
        st->frame = frame;
        st->n_fluid = uint32_t(N);
        st->n_boundary = uint32_t(B);
        if (cfg.zsort_interval && (frame % cfg.zsort_interval) == 1u) {
            zsort();
            st->zsorted = 1;
        }
        for (int32_t ph = MPS_PHASE_HASH; ph <= MPS_PHASE_ADVECT; ph++) {
            const int rc = run_phase(ph, st);
            if (rc) return rc;
        }
        st->nnz_ff = ff.nei.size();
        st->nnz_fb = fb.nei.size();
        st->nnz_bb = bb.nei.size();
        st->nonfinite = finite_state() ? 0u : 1u;
        frame++;
        return MPS_OK;
    }

    // ---------------------------------------------------------------- array access
    struct Ref {
        void* p;
        size_t count, eb;
    };
    Ref ref(int32_t id) {
        auto R = [](auto& vec, size_t eb) { return Ref{(void*)vec.data(), vec.size() * sizeof(vec[0]) / eb, eb}; };
        const size_t t = sizeof(T);
        switch (id) {
            case MPS_ARR_FLUID_POS: return R(x, 3 * t);
            case MPS_ARR_FLUID_VEL: return R(v, 3 * t);
            case MPS_ARR_FLUID_RADIUS: return R(r, t);
            case MPS_ARR_FLUID_MASS: return R(m, t);
            case MPS_ARR_FLUID_COLOR: return R(color, 16);
            case MPS_ARR_FLUID_DENSITY: return R(rho, t);
            case MPS_ARR_FLUID_ALPHA: return R(alpha, t);
            case MPS_ARR_FLUID_KAPPA: return R(kappa, t);
            case MPS_ARR_FLUID_DELTA: return R(delta, t);
            case MPS_ARR_FLUID_FACTOR_ST: return R(fst, t);
            case MPS_ARR_FLUID_CG_B: return R(cgb, 3 * t);
            case MPS_ARR_BND_POS: return R(bx, 3 * t);
            case MPS_ARR_BND_RADIUS: return R(br, t);
            case MPS_ARR_BND_VOLUME: return R(bvol, t);
            case MPS_ARR_BND_MASS: return R(bm, t);
            case MPS_ARR_BND_VEL: return R(bv, 3 * t);
            case MPS_ARR_BND_FACE_ID: return R(bface, 4);
            case MPS_ARR_BND_BCC: return R(bbcc, 2 * t);
            case MPS_ARR_BND_OBJECT: return R(bobj, 4);
            case MPS_ARR_FHASH_KEY: return R(fh.key, 4);
            case MPS_ARR_FHASH_ID: return R(fh.id, 4);
            case MPS_ARR_FHASH_START: return R(fh.start, 4);
            case MPS_ARR_FHASH_END: return R(fh.end, 4);
            case MPS_ARR_BHASH_KEY: return R(bh.key, 4);
            case MPS_ARR_BHASH_ID: return R(bh.id, 4);
            case MPS_ARR_BHASH_START: return R(bh.start, 4);
            case MPS_ARR_BHASH_END: return R(bh.end, 4);
            case MPS_ARR_NEI_FF_IDX: return R(ff.idx, 4);
            case MPS_ARR_NEI_FF: return R(ff.nei, 4);
            case MPS_ARR_NEI_FB_IDX: return R(fb.idx, 4);
            case MPS_ARR_NEI_FB: return R(fb.nei, 4);
            case MPS_ARR_NEI_BB_IDX: return R(bb.idx, 4);
            case MPS_ARR_NEI_BB: return R(bb.nei, 4);
            default: return Ref{nullptr, 0, 0};
        }
    }
    int array_info(int32_t id, size_t* count, size_t* eb) override {
        const Ref q = ref(id);
        if (q.eb == 0) { err = "unknown array id"; return MPS_ERR_INVALID; }
        if (count) *count = q.count;
        if (eb) *eb = q.eb;
        return MPS_OK;
    }
    int get_array(int32_t id, void* dst, size_t bytes) override {
        const Ref q = ref(id);
        if (q.eb == 0) { err = "unknown array id"; return MPS_ERR_INVALID; }
        if (bytes != q.count * q.eb) { err = "size mismatch"; return MPS_ERR_INVALID; }
        if (bytes) std::memcpy(dst, q.p, bytes);
        return MPS_OK;
    }
    int set_array(int32_t id, const void* src, size_t bytes) override {
        const Ref q = ref(id);
        if (q.eb == 0) { err = "unknown array id"; return MPS_ERR_INVALID; }
        if (bytes != q.count * q.eb) { err = "size mismatch"; return MPS_ERR_INVALID; }
        if (bytes) std::memcpy(q.p, src, bytes);
        return MPS_OK;
    }
};

}  // namespace

struct orc_ctx {
    std::unique_ptr<Base> impl;
};

// A coupled world of fluid objects ("phases"), each a complete context with its own material,
// hash, lists and walls. Phase k's ref lists point into every other phase in ascending order
// (the reference's ref overloads, which its own step never calls). The step runs OnUpdate's
// order over all phases (MPSSystem.cpp:191-291); the pressure loops run jointly: one
// evaluation of every phase per iteration, each phase's error E_k = sum of its positive
// errors / its particle count as in ComputeDFSPHDivergenceFree / ConstantDensity
// (MPSSPHDFSPHUtil.cu:270-441), the loop ending at the first iteration where every phase has
// E_k < tol and l > l_min; then every phase applies. Viscosity and surface tension stay within
// each phase (the reference's CG takes the object's own list only).
template <class T>
struct World {
    std::vector<Ctx<T>*> P;
    void wire() {
        for (size_t k = 0; k < P.size(); k++) {
            if (P[k]->refs.size() == P.size() - 1) continue;
            P[k]->refs.clear();
            for (size_t t = 0; t < P.size(); t++)
                if (t != k) P[k]->refs.push_back(typename Ctx<T>::RefPhase{P[t], CSR{}});
        }
    }
    void pressure_joint(bool df, mps_step_stats* st) {
        const mps_config& cfg = P[0]->cfg;
        const double tol = df ? cfg.df_tol : cfg.cd_tol;
        const uint32_t lmin = df ? cfg.df_min_iter : cfg.cd_min_iter;
        uint32_t l = 1u;
        bool capped = true;
        while (l <= cfg.max_iter) {
            for (Ctx<T>* p : P) p->density_delta();
            bool ok = true;
            for (size_t k = 0; k < P.size(); k++) {
                const double E = P[k]->N ? P[k]->stiffness(df) / double(P[k]->N) : 0.0;
                double* log = df ? st[k].err_df : st[k].err_cd;
                if (l - 1 < MPS_MAX_ITER_LOG) log[l - 1] = E;
                ok = ok && E < tol && l > lmin;
            }
            if (ok) {
                capped = false;
                break;
            }
            for (Ctx<T>* p : P) p->apply_pressure();
            l++;
        }
        for (size_t k = 0; k < P.size(); k++) {
            if (df) { st[k].l_df = l; st[k].hit_cap_df = capped; }
            else { st[k].l_cd = l; st[k].hit_cap_cd = capped; }
        }
    }
    int run_phase(int32_t ph, mps_step_stats* st) {
        wire();
        switch (ph) {
            case MPS_PHASE_NEIGHBORS:
                for (Ctx<T>* p : P) p->neighbors();
                for (Ctx<T>* p : P) p->build_ref_lists();
                return MPS_OK;
            case MPS_PHASE_DIVERGENCE: pressure_joint(true, st); return MPS_OK;
            case MPS_PHASE_CONSTANT_DENSITY: pressure_joint(false, st); return MPS_OK;
            default:
                for (size_t k = 0; k < P.size(); k++) {
                    const int rc = P[k]->run_phase(ph, &st[k]);
                    if (rc) return rc;
                }
                return MPS_OK;
        }
    }
    int step(mps_step_stats* st) {
        const uint32_t frame = P[0]->frame;
        const mps_config& cfg = P[0]->cfg;
        for (size_t k = 0; k < P.size(); k++) {
            std::memset(&st[k], 0, sizeof(st[k]));
            st[k].frame = frame;
            st[k].n_fluid = uint32_t(P[k]->N);
            st[k].n_boundary = uint32_t(P[k]->B);
        }
        if (cfg.zsort_interval && (frame % cfg.zsort_interval) == 1u)
            for (size_t k = 0; k < P.size(); k++) {
                P[k]->zsort();
                st[k].zsorted = 1;
            }
        for (int32_t ph = MPS_PHASE_HASH; ph <= MPS_PHASE_ADVECT; ph++) {
            const int rc = run_phase(ph, st);
            if (rc) return rc;
        }
        for (size_t k = 0; k < P.size(); k++) {
            Ctx<T>* p = P[k];
            st[k].nnz_ff = p->ff.nei.size();
            st[k].nnz_fb = p->fb.nei.size();
            st[k].nnz_bb = p->bb.nei.size();
            st[k].nonfinite = p->finite_state() ? 0u : 1u;
            p->frame++;
        }
        return MPS_OK;
    }
};

struct orc_world {
    int precision = 32;
    std::vector<orc_ctx*> ph;
    std::string err;
    template <class T>
    World<T> world() {
        World<T> w;
        for (orc_ctx* c : ph) w.P.push_back(static_cast<Ctx<T>*>(c->impl.get()));
        return w;
    }
};

extern "C" {

int orc_create(const mps_config* cfg, int precision, orc_ctx** out) {
    if (!cfg || !out || (precision != 32 && precision != 64)) return MPS_ERR_INVALID;
    for (int k = 0; k < 3; k++)
        if (cfg->hash_size[k] == 0 || (cfg->hash_size[k] & (cfg->hash_size[k] - 1))) return MPS_ERR_INVALID;
    auto* c = new orc_ctx;
    if (precision == 64) {
        auto p = std::make_unique<Ctx<double>>();
        p->cfg = *cfg;
        p->cubic = cfg->kernel == MPS_KERNEL_CUBIC;
        c->impl = std::move(p);
    } else {
        auto p = std::make_unique<Ctx<float>>();
        p->cfg = *cfg;
        p->cubic = cfg->kernel == MPS_KERNEL_CUBIC;
        c->impl = std::move(p);
    }
    *out = c;
    return MPS_OK;
}
void orc_destroy(orc_ctx* c) { delete c; }

int orc_set_boundary_objects(orc_ctx* c, uint32_t k, const uint32_t* counts) {
    return c ? c->impl->boundary_objects(k, counts) : MPS_ERR_INVALID;
}

int orc_world_create(const mps_config* cfg, int precision, uint32_t nphase, orc_world** out) {
    if (!cfg || !out || nphase == 0) return MPS_ERR_INVALID;
    auto* w = new orc_world;
    w->precision = precision;
    for (uint32_t k = 0; k < nphase; k++) {
        orc_ctx* c = nullptr;
        const int rc = orc_create(cfg, precision, &c);
        if (rc) {
            for (orc_ctx* q : w->ph) delete q;
            delete w;
            return rc;
        }
        w->ph.push_back(c);
    }
    *out = w;
    return MPS_OK;
}
void orc_world_destroy(orc_world* w) {
    if (!w) return;
    for (orc_ctx* c : w->ph) delete c;
    delete w;
}
orc_ctx* orc_world_phase(orc_world* w, uint32_t k) { return w && k < w->ph.size() ? w->ph[k] : nullptr; }
int orc_world_step(orc_world* w, mps_step_stats* st) {
    if (!w || !st) return MPS_ERR_INVALID;
    if (w->precision == 64) {
        World<double> x = w->world<double>();
        return x.step(st);
    }
    World<float> x = w->world<float>();
    return x.step(st);
}
int orc_world_run_phase(orc_world* w, int32_t phase, mps_step_stats* st) {
    if (!w || !st) return MPS_ERR_INVALID;
    if (w->precision == 64) {
        World<double> x = w->world<double>();
        return x.run_phase(phase, st);
    }
    World<float> x = w->world<float>();
    return x.run_phase(phase, st);
}
const char* orc_last_error(const orc_ctx* c) { return c ? c->impl->err.c_str() : "null context"; }
int orc_precision(const orc_ctx* c) { return c ? c->impl->precision() : 0; }
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
int orc_set_physics(orc_ctx* c, const mps_physics* p) { return (c && p) ? c->impl->set_physics(p) : MPS_ERR_INVALID; }
int orc_set_fluid_material(orc_ctx* c, const mps_sph_material* m) {
    return (c && m) ? c->impl->set_fluid_material(m) : MPS_ERR_INVALID;
}
int orc_upload_fluid(orc_ctx* c, size_t n, const double* pos, const double* vel, const double* radius,
                     const double* mass) {
    if (!c || (n && (!pos || !radius))) return MPS_ERR_INVALID;
    return c->impl->upload_fluid(n, pos, vel, radius, mass);
}
int orc_upload_boundary(orc_ctx* c, size_t n, const double* pos, const double* radius, const double* mass) {
    if (!c || (n && (!pos || !radius))) return MPS_ERR_INVALID;
    return c->impl->upload_boundary(n, pos, radius, mass);
}
int orc_sample_boundary_mesh(orc_ctx* c, const double* verts, size_t nv, const uint32_t* tris, size_t nt,
                             const double* center, const double* size, const mps_mesh_material* m) {
    if (!c || !m) return MPS_ERR_INVALID;
    return c->impl->sample_mesh(verts, nv, tris, nt, center, size, m);
}
int orc_resample_boundary_mesh(orc_ctx* c, const double* verts, size_t nv, const mps_mesh_material* m,
                               size_t* regenerated) {
    if (!c || !m) return MPS_ERR_INVALID;
    return c->impl->resample_mesh(verts, nv, m, regenerated);
}
int orc_step(orc_ctx* c, mps_step_stats* out) { return c ? c->impl->step(out) : MPS_ERR_INVALID; }
int orc_run_phase(orc_ctx* c, int32_t phase, mps_step_stats* out) {
    return c ? c->impl->run_phase(phase, out) : MPS_ERR_INVALID;
}
int orc_array_info(orc_ctx* c, int32_t id, size_t* count, size_t* eb) {
    return c ? c->impl->array_info(id, count, eb) : MPS_ERR_INVALID;
}
int orc_get_array(orc_ctx* c, int32_t id, void* dst, size_t bytes) {
    return c ? c->impl->get_array(id, dst, bytes) : MPS_ERR_INVALID;
}
int orc_set_array(orc_ctx* c, int32_t id, const void* src, size_t bytes) {
    return c ? c->impl->set_array(id, src, bytes) : MPS_ERR_INVALID;
}
int orc_get_frame(orc_ctx* c, uint32_t* f) {
    if (!c || !f) return MPS_ERR_INVALID;
    *f = c->impl->frame;
    return MPS_OK;
}
int orc_set_frame(orc_ctx* c, uint32_t f) {
    if (!c) return MPS_ERR_INVALID;
    c->impl->frame = f;
    return MPS_OK;
}

int orc_cg_operator(orc_ctx* c, const void* p, void* out, double f) {
    if (!c || !p || !out) return MPS_ERR_INVALID;
    return c->impl->cg_operator(p, out, f);
}

int orc_kernel_eval(int precision, int kernel, size_t n, const double* dist, double invh, double* w, double* g,
                    double* stw, double* stg) {
    const int cubic = kernel == MPS_KERNEL_CUBIC;
    for (size_t i = 0; i < n; i++) {
        if (precision == 64) {
            if (w) w[i] = Wk<double>(dist[i], invh, cubic);
            if (g) g[i] = Gk<double>(dist[i], invh, cubic);
            if (stw) stw[i] = STWk<double>(dist[i], invh);
            if (stg) stg[i] = STGk<double>(dist[i], invh);
        } else {
            const float d = float(dist[i]), ih = float(invh);
            if (w) w[i] = Wk<float>(d, ih, cubic);
            if (g) g[i] = Gk<float>(d, ih, cubic);
            if (stw) stw[i] = STWk<float>(d, ih);
            if (stg) stg[i] = STGk<float>(d, ih);
        }
    }
    return MPS_OK;
}
int orc_cell_keys(int precision, size_t n, const double* pos, double cell, const uint32_t* H, int morton,
                  uint32_t* keys) {
    const double c = precision == 64 ? cell : double(float(cell));
    for (size_t i = 0; i < n; i++) {
        double p[3];
        for (int k = 0; k < 3; k++) p[k] = precision == 64 ? pos[3 * i + k] : double(float(pos[3 * i + k]));
        const IV3 g = grid_pos(p[0], p[1], p[2], c, H);
        keys[i] = morton ? grid_zindex(g, H) : grid_index(g, H);
    }
    return MPS_OK;
}
int orc_mesh_rtri(size_t nv, const uint32_t* tris, size_t nt, uint32_t* rtri) {
    rtri_build(nv, tris, nt, rtri);
    return MPS_OK;
}
int orc_mesh_sample_counts(const double* verts, size_t nv, const uint32_t* tris, size_t nt, double radius,
                           uint32_t* counts) {
    Mesh M;
    M.v.assign(verts, verts + 3 * nv);
    M.f.assign(tris, tris + 3 * nt);
    M.rtri.resize(nt);
    rtri_build(nv, tris, nt, M.rtri.data());
    for (size_t t = 0; t < nt; t++) sample_face(M, t, radius, nullptr, &counts[t]);
    return MPS_OK;
}

}  // extern "C"
