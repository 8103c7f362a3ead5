/* mps_system.hpp — the reference's host classes, by name, over the C ABI of mps_sph.h.
 *
 * A program written against MPS_CUDA's System / SPHMaterial / MeshMaterial / MeshObject /
 * SpatialHash / kernel::ParticleSampling keeps its calls; the state they set up lives in one
 * mps_ctx (device memory and a HIP stream) owned by mps::System. Header-only C++17; link
 * with libmps_sph.so. Errors throw mps::Exception carrying the library's status code and
 * message (the reference exits or asserts instead).
 *
 *   mps::System                    MPSSystem.h:23-76, MPSSystem.cpp:19-291,357-651
 *   mps::SPHMaterial               MPSSPHMaterial.h:9-43, MPSSPHMaterial.cpp:4-44
 *   mps::MeshMaterial              MPSMeshMaterial.h:9-28, MPSMeshMaterial.cpp:4-20
 *   mps::PhysicsParam              MPSUniformDef.h:19-25
 *   mps::SPHObject                 MPSSPHObject.h:10-38 (host arrays, uploaded on commit)
 *   mps::BoundaryParticleObject    MPSBoundaryParticleObject.h:10-38 (the sampled walls)
 *   mps::MeshObject::LoadMesh      MPSMeshObject.cu:58-103
 *   mps::SpatialHash               MPSSpatialHash.h:20-57 (sizes; UpdateHash / ZSort /
 *                                  BuildNeighorhood run the context's phases)
 *   mps::kernel::ParticleSampling  MPSParticleSamplingUtil.h:13-14, .cu:14-133 (a second call
 *                                  after MeshObject::SetVertexPositions re-samples)
 */
#ifndef MPS_SYSTEM_HPP
#define MPS_SYSTEM_HPP

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mps_sph.h"

namespace mps {

using REAL = double;  // MPSDef.h:5-28 (the library computes in float; parameters stay double)

struct REAL3 {
    REAL x = 0.0, y = 0.0, z = 0.0;
};
inline REAL3 operator+(REAL3 a, REAL3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline REAL3 operator-(REAL3 a, REAL3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline REAL3 operator*(REAL3 a, REAL s) { return {a.x * s, a.y * s, a.z * s}; }
inline bool operator==(REAL3 a, REAL3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }

struct uvec3 {
    uint32_t x = 0, y = 0, z = 0;
};

class Exception : public std::runtime_error {
public:
    Exception(int code, const std::string& what) : std::runtime_error(what), m_code(code) {}
    int code() const { return m_code; }

private:
    int m_code;
};

namespace detail {
inline void check(const mps_ctx* c, int rc, const char* what) {
    if (rc == MPS_OK) return;
    const char* msg = c ? mps_last_error(c) : nullptr;
    throw Exception(rc, std::string(what) + " failed (status " + std::to_string(rc) + ")" +
                            (msg && *msg ? std::string(": ") + msg : std::string()));
}
}  // namespace detail

// PhysicsParam (MPSUniformDef.h:19-25)
struct PhysicsParam {
    REAL dt = 0.0;
    REAL3 gravity;
    REAL3 min, max;
};

// SPHMaterialParam / SPHMaterial (MPSSPHMaterial.h:9-43, .cpp:4-44)
struct SPHMaterialParam {
    REAL radius, mass, density, volume, viscosity, XSPH, surfaceTension;
    std::array<float, 4> color;
};
class SPHMaterial {
public:
    SPHMaterial() {
        SetParam(0.01, 1.0);
        SetViscosity(0.025);
        SetXSPH(0.1);
        SetSurfaceTension(100.0);
        SetColor({0.3f, 0.8f, 0.2f, 1.0f});
    }
    void SetParam(REAL radius, REAL density) {
        m_param.radius = radius;
        m_param.volume = radius * radius * radius / 48.0 * 3.141592;
        m_param.density = density;
        m_param.mass = density * m_param.volume;
    }
    void SetRadius(REAL radius) { SetParam(radius, GetDensity()); }
    void SetViscosity(REAL viscosity) { m_param.viscosity = viscosity; }
    void SetXSPH(REAL XSPH) { m_param.XSPH = XSPH; }
    void SetSurfaceTension(REAL surfaceTension) { m_param.surfaceTension = surfaceTension; }
    void SetColor(const std::array<float, 4>& color) { m_param.color = color; }
    REAL GetRadius() const { return m_param.radius; }
    REAL GetMass() const { return m_param.mass; }
    REAL GetDensity() const { return m_param.density; }
    REAL GetVolume() const { return m_param.volume; }
    REAL GetViscosity() const { return m_param.viscosity; }
    REAL GetXSPH() const { return m_param.XSPH; }
    REAL GetSurfaceTension() const { return m_param.surfaceTension; }
    const std::array<float, 4>& GetColor() const { return m_param.color; }
    const SPHMaterialParam& GetParam() const { return m_param; }
    mps_sph_material ToABI() const {
        return mps_sph_material{m_param.radius, m_param.density, m_param.viscosity, m_param.XSPH, m_param.surfaceTension};
    }

private:
    SPHMaterialParam m_param{};
};

// MeshMaterialParam / MeshMaterial (MPSMeshMaterial.h:9-28, .cpp:4-20)
struct MeshMaterialParam {
    REAL radius, mass, density, viscosity, surfaceTension;
    std::array<float, 4> frontColor, backColor;
};
class MeshMaterial {
public:
    MeshMaterial() {
        SetParam(0.1, 1.0);
        m_param.viscosity = 0.0;
        m_param.surfaceTension = 0.0;
        m_param.frontColor = {0.0f, 0.0f, 0.0f, 1.0f};
        m_param.backColor = {0.0f, 0.0f, 0.0f, 1.0f};
    }
    void SetParam(REAL radius, REAL density) {
        m_param.radius = radius;
        const REAL volume = radius * radius * radius / 48.0 * 3.141592;
        m_param.density = density;
        m_param.mass = density * volume;
    }
    const MeshMaterialParam& GetParam() const { return m_param; }

private:
    MeshMaterialParam m_param{};
};

// Fluid particles as the host hands them over (SPHObject's arrays before CopyFromHost).
class SPHObject {
public:
    void Resize(size_t size) {
        m_position.resize(size);
        m_radius.resize(size);
        m_velocity.resize(size);
        m_mass.resize(size);
        m_color.resize(size);
    }
    void Clear() { Resize(0); }
    size_t GetSize() const { return m_radius.size(); }

    std::vector<REAL3> m_position;
    std::vector<REAL> m_radius;
    std::vector<REAL3> m_velocity;
    std::vector<REAL> m_mass;
    std::vector<std::array<float, 4>> m_color;
};

// The triangle mesh of the walls: LoadMesh keeps the file's vertices and faces plus the box
// it is fitted into (Translate) and the face density (InitFaceInfo); ParticleSampling applies
// them on the device.
class MeshObject {
public:
    void LoadMesh(const std::string& filePath, const REAL3& vCenter, const REAL3& vSize, REAL density) {
        size_t nv = 0, nt = 0;
        detail::check(nullptr, mps_read_obj(nullptr, filePath.c_str(), nullptr, &nv, nullptr, &nt), "MeshObject::LoadMesh");
        m_vertices.assign(3 * nv, 0.0);
        m_faces.assign(3 * nt, 0u);
        detail::check(nullptr, mps_read_obj(nullptr, filePath.c_str(), m_vertices.data(), &nv, m_faces.data(), &nt),
                      "MeshObject::LoadMesh");
        m_center = vCenter;
        m_size = vSize;
        m_density = density;
        m_world.clear();
    }
    // Moves the vertices (world coordinates, after the fit; same count): the next
    // ParticleSampling into a boundary already sampled from this mesh re-samples only the
    // faces whose sampling changed (MPSParticleSamplingUtil.cu:14-133).
    void SetVertexPositions(const std::vector<double>& world) {
        if (world.size() != m_vertices.size()) throw Exception(MPS_ERR_INVALID, "MeshObject::SetVertexPositions: vertex count");
        m_world = world;
    }
    const std::vector<double>& GetVertexPositions() const { return m_world; }
    size_t GetSize() const { return m_faces.size() / 3; }
    size_t GetVertexSize() const { return m_vertices.size() / 3; }
    const std::vector<double>& GetVertices() const { return m_vertices; }
    const std::vector<uint32_t>& GetFaces() const { return m_faces; }
    const REAL3& GetCenter() const { return m_center; }
    const REAL3& GetFitSize() const { return m_size; }
    REAL GetDensity() const { return m_density; }

private:
    std::vector<double> m_vertices, m_world;
    std::vector<uint32_t> m_faces;
    REAL3 m_center, m_size;
    REAL m_density = 1.0;
};

// The sampled wall particles; they live in the context of the System that owns the object.
class BoundaryParticleObject {
public:
    explicit BoundaryParticleObject(mps_ctx* const* ctx = nullptr) : m_ctx(ctx) {}
    size_t GetSize() const {
        if (!m_ctx || !*m_ctx) return 0;
        size_t n = 0, e = 0;
        detail::check(*m_ctx, mps_array_info(*m_ctx, MPS_ARR_BND_RADIUS, &n, &e), "BoundaryParticleObject::GetSize");
        return n;
    }
    mps_ctx* context() const { return m_ctx ? *m_ctx : nullptr; }
    // the mesh and context of the last full sampling (a later call with moved vertices re-samples)
    const void* m_sampledFrom = nullptr;
    const mps_ctx* m_sampledIn = nullptr;

private:
    mps_ctx* const* m_ctx;
};

// Hash sizes of one particle set (SpatialHash::SetObjectSize / SetCeilSize / SetHashSize).
// UpdateHash, ZSort and BuildNeighorhood run the context's phases, which treat the walls and
// the fluid together in OnUpdate's order; the lists are read back with the getters.
class SpatialHash {
public:
    explicit SpatialHash(mps_ctx* const* ctx = nullptr, bool fluid = true) : m_ctx(ctx), m_fluid(fluid) {}
    void SetObjectSize(size_t size) { m_objSize = size; }
    void SetCeilSize(REAL size) { m_ceilSize = size; }
    void SetHashSize(const uvec3& size) { m_hashSize = size; }
    size_t GetObjectSize() const { return m_objSize; }
    REAL GetCeilSize() const { return m_ceilSize; }
    const uvec3& GetHashSize() const { return m_hashSize; }

    void UpdateHash() { phase(MPS_PHASE_HASH, "SpatialHash::UpdateHash"); }
    void ZSort() { phase(MPS_PHASE_ZSORT, "SpatialHash::ZSort"); }
    void BuildNeighorhood() { phase(MPS_PHASE_NEIGHBORS, "SpatialHash::BuildNeighorhood"); }

    std::vector<uint32_t> GetHashKeys() const { return get(m_fluid ? MPS_ARR_FHASH_KEY : MPS_ARR_BHASH_KEY); }
    std::vector<uint32_t> GetHashIDs() const { return get(m_fluid ? MPS_ARR_FHASH_ID : MPS_ARR_BHASH_ID); }
    std::vector<uint32_t> GetStartIdx() const { return get(m_fluid ? MPS_ARR_FHASH_START : MPS_ARR_BHASH_START); }
    std::vector<uint32_t> GetEndIdx() const { return get(m_fluid ? MPS_ARR_FHASH_END : MPS_ARR_BHASH_END); }
    // GetNeighborhood (MPSSpatialHash.cu:44-51) as CSR copies: (offsets, ids) of the set's
    // own list, or with refBoundary the fluid->wall list.
    std::pair<std::vector<uint32_t>, std::vector<uint32_t>> GetNeighborhood(bool refBoundary = false) const {
        if (!m_fluid) return {get(MPS_ARR_NEI_BB_IDX), get(MPS_ARR_NEI_BB)};
        if (refBoundary) return {get(MPS_ARR_NEI_FB_IDX), get(MPS_ARR_NEI_FB)};
        return {get(MPS_ARR_NEI_FF_IDX), get(MPS_ARR_NEI_FF)};
    }

private:
    mps_ctx* ctx(const char* what) const {
        if (!m_ctx || !*m_ctx) throw Exception(MPS_ERR_STATE, std::string(what) + ": the System is not initialised");
        return *m_ctx;
    }
    void phase(int32_t ph, const char* what) {
        mps_ctx* c = ctx(what);
        detail::check(c, mps_run_phase(c, ph, nullptr), what);
    }
    std::vector<uint32_t> get(int32_t id) const {
        mps_ctx* c = ctx("SpatialHash");
        size_t n = 0, e = 0;
        detail::check(c, mps_array_info(c, id, &n, &e), "SpatialHash");
        std::vector<uint32_t> out(n);
        detail::check(c, mps_get_array(c, id, out.data(), n * e), "SpatialHash");
        return out;
    }
    mps_ctx* const* m_ctx;
    bool m_fluid;
    size_t m_objSize = 0;
    REAL m_ceilSize = 0.0;
    uvec3 m_hashSize{64, 64, 64};
};

namespace kernel {
namespace ParticleSampling {
// ParticleSampling(material, mesh, boundary) (MPSParticleSamplingUtil.cu:14-133) after
// LoadMesh's Translate and InitFaceInfo: samples the mesh into the boundary's context.
// Called again for the same mesh and boundary after MeshObject::SetVertexPositions, it
// re-samples the faces whose sample count or shortest edge changed and keeps the others.
inline void ParticleSampling(const MeshMaterialParam& material, const MeshObject& mesh, BoundaryParticleObject& boundary) {
    mps_ctx* c = boundary.context();
    if (!c) throw Exception(MPS_ERR_STATE, "ParticleSampling: the boundary object has no context");
    if (boundary.m_sampledFrom == &mesh && boundary.m_sampledIn == c && !mesh.GetVertexPositions().empty()) {
        const mps_mesh_material m{material.radius, mesh.GetDensity()};
        detail::check(c,
                      mps_resample_boundary_mesh(c, mesh.GetVertexPositions().data(), mesh.GetVertexSize(), &m, nullptr),
                      "ParticleSampling");
        return;
    }
    const double center[3] = {mesh.GetCenter().x, mesh.GetCenter().y, mesh.GetCenter().z};
    const double size[3] = {mesh.GetFitSize().x, mesh.GetFitSize().y, mesh.GetFitSize().z};
    const mps_mesh_material m{material.radius, mesh.GetDensity()};
    detail::check(c,
                  mps_sample_boundary_mesh(c, mesh.GetVertices().data(), mesh.GetVertexSize(), mesh.GetFaces().data(),
                                           mesh.GetSize(), center, size, &m),
                  "ParticleSampling");
    boundary.m_sampledFrom = &mesh;
    boundary.m_sampledIn = c;
}
}  // namespace ParticleSampling
}  // namespace kernel

// System (MPSSystem.h:23-76): the simulation part of the reference's System. Initalize builds
// SurfaceTensionTestScene (MPSSystem.cpp:552-651) around a wall mesh file (the reference
// loads obj/cube.obj); OnUpdate is one step when running; OnKeyDown handles the simulation
// keys (' ' run/pause, '[' / ']' resize, 'G' gravity, MPSSystem.cpp:115-146).
class System {
public:
    System() : m_boundary(&m_ctx), m_sphHash(&m_ctx, true), m_boundaryHash(&m_ctx, false) { mps_config_init(&m_cfg); }
    ~System() { Release(); }
    System(const System&) = delete;
    System& operator=(const System&) = delete;

    void SetDevice(int device) { m_cfg.device = device; }
    // Configuration of the context created by Initalize / Commit (tolerances, kernel, ...).
    mps_config& Config() { return m_cfg; }

    void Initalize(const std::string& meshPath) {
        m_meshPath = meshPath;
        m_physicsParam.gravity = {0.0, -9.8, 0.0};
        m_physicsParam.dt = 0.005;
        m_physicsParam.min = {-5.0, -5.0, -5.0};
        m_physicsParam.max = {5.0, 5.0, 5.0};
        m_sphMaterial.SetRadius(kRadius * 4);
        m_particleWSize = 5;
        ResizeParticle();
    }

    void OnUpdate() {
        if (!b_runSim) return;
        Update();
    }
    // One step regardless of the run flag.
    void Update() {
        if (!m_ctx) throw Exception(MPS_ERR_STATE, "System::Update: not initialised");
        if (m_physicsDirty) {
            const mps_physics p = physicsABI();
            detail::check(m_ctx, mps_set_physics(m_ctx, &p), "mps_set_physics");
            m_physicsDirty = false;
        }
        detail::check(m_ctx, mps_step(m_ctx, &m_stats), "System::OnUpdate");
        m_frame++;
    }

    void OnKeyDown(uint32_t key) {
        const uint32_t offset = 5;
        switch (key) {
            case 32: b_runSim = !b_runSim; break;
            case 219:
                m_particleWSize = m_particleWSize > offset ? m_particleWSize - offset : 1;
                ResizeParticle();
                break;
            case 221:
                m_particleWSize = m_particleWSize + offset;
                ResizeParticle();
                break;
            case 'g':
            case 'G':
                m_physicsParam.gravity = m_physicsParam.gravity == REAL3{} ? REAL3{0.0, -9.8, 0.0} : REAL3{};
                m_physicsDirty = true;
                break;
            default: break;
        }
    }

    // Fluid, walls and hash sizes as set on the objects below go into a fresh context.
    void Commit() {
        Release();
        m_cfg.hash_size[0] = m_sphHash.GetHashSize().x;
        m_cfg.hash_size[1] = m_sphHash.GetHashSize().y;
        m_cfg.hash_size[2] = m_sphHash.GetHashSize().z;
        m_cfg.cell_size = m_sphHash.GetCeilSize();
        detail::check(nullptr, mps_create(&m_cfg, &m_ctx), "mps_create");
        const mps_physics p = physicsABI();
        detail::check(m_ctx, mps_set_physics(m_ctx, &p), "mps_set_physics");
        const mps_sph_material m = m_sphMaterial.ToABI();
        detail::check(m_ctx, mps_set_fluid_material(m_ctx, &m), "mps_set_fluid_material");
        if (m_mesh.GetSize()) kernel::ParticleSampling::ParticleSampling(m_meshMaterial.GetParam(), m_mesh, m_boundary);
        const size_t n = m_sph.GetSize();
        std::vector<float> pos(3 * n), vel(3 * n), rad(n), mass(n);
        for (size_t i = 0; i < n; i++) {
            pos[3 * i] = float(m_sph.m_position[i].x);
            pos[3 * i + 1] = float(m_sph.m_position[i].y);
            pos[3 * i + 2] = float(m_sph.m_position[i].z);
            vel[3 * i] = float(m_sph.m_velocity[i].x);
            vel[3 * i + 1] = float(m_sph.m_velocity[i].y);
            vel[3 * i + 2] = float(m_sph.m_velocity[i].z);
            rad[i] = float(m_sph.m_radius[i]);
            mass[i] = float(m_sph.m_mass[i]);
        }
        detail::check(m_ctx, mps_upload_fluid(m_ctx, n, pos.data(), vel.data(), rad.data(), mass.data()), "mps_upload_fluid");
        m_physicsDirty = false;
        m_frame = 0;
    }

    void Release() {
        if (m_ctx) mps_destroy(m_ctx);
        m_ctx = nullptr;
    }

    // Positions (x, y, z per particle, the caller's order) of the current state.
    std::vector<float> GetPositions() const {
        if (!m_ctx) return {};
        size_t n = 0, e = 0;
        detail::check(m_ctx, mps_array_info(m_ctx, MPS_ARR_FLUID_POS, &n, &e), "GetPositions");
        std::vector<float> out(3 * n);
        detail::check(m_ctx, mps_get_array(m_ctx, MPS_ARR_FLUID_POS, out.data(), n * e), "GetPositions");
        return out;
    }

    bool IsRunning() const { return b_runSim; }
    uint64_t GetFrame() const { return m_frame; }
    const mps_step_stats& GetLastStats() const { return m_stats; }
    uint32_t GetParticleWSize() const { return m_particleWSize; }
    PhysicsParam& GetPhysicsParam() { m_physicsDirty = true; return m_physicsParam; }
    SPHMaterial& GetSPHMaterial() { return m_sphMaterial; }
    MeshMaterial& GetMeshMaterial() { return m_meshMaterial; }
    SPHObject& GetSPHObject() { return m_sph; }
    MeshObject& GetMeshObject() { return m_mesh; }
    BoundaryParticleObject& GetBoundaryParticleObject() { return m_boundary; }
    SpatialHash& GetSPHHash() { return m_sphHash; }
    SpatialHash& GetBoundaryParticleHash() { return m_boundaryHash; }
    mps_ctx* GetContext() const { return m_ctx; }

private:
    static constexpr REAL kRadius = 0.1;  // MPSSystem.cpp:14-17

    void ResizeParticle() { SurfaceTensionTestScene(); }

    // MPSSystem.cpp:552-651: two size^3 blocks side by side in a box 3.5 wider than them,
    // inside walls sampled from the mesh fitted to that box.
    void SurfaceTensionTestScene() {
        const uint32_t size = m_particleWSize;
        const size_t nSize = size_t(size) * size * size * 2;
        m_sphMaterial.SetRadius(kRadius * 4);
        const REAL gap = kRadius * 0.9f;
        const REAL offset = 0.001f;
        const REAL stride = gap + gap + offset;
        const REAL l = stride * (size - 1) + gap + gap;
        const REAL startPos = -l;
        m_physicsParam.min = {startPos - l * 0.025 + gap - 3.5, startPos + gap - 3.5, startPos + gap - 3.5};
        m_physicsParam.max = {l * 0.025 + (size - 1) * stride + gap + 3.5, startPos + (size - 1) * stride + gap + 3.5,
                              startPos + (size - 1) * stride + gap + 3.5};
        m_sph.Resize(0);
        m_sph.m_position.reserve(nSize);
        for (uint32_t z = 0; z < size; z++)
            for (uint32_t y = 0; y < size; y++)
                for (uint32_t x = 0; x < size; x++) {
                    m_sph.m_position.push_back(REAL3{startPos - l * 0.025 + x * stride + gap, startPos + y * stride + gap,
                                                     startPos + z * stride + gap});
                    m_sph.m_position.push_back(
                        REAL3{l * 0.025 + x * stride + gap, startPos + y * stride + gap, startPos + z * stride + gap});
                }
        m_sph.m_radius.assign(nSize, kRadius);
        m_sph.m_velocity.assign(nSize, REAL3{});
        m_sph.m_mass.assign(nSize, m_sphMaterial.GetMass());
        m_sph.m_color.assign(nSize, {0.0f, 0.0f, 1.0f, 1.0f});
        m_sphHash.SetObjectSize(nSize);
        m_sphHash.SetCeilSize(kRadius * 4);
        m_sphHash.SetHashSize({64, 64, 64});

        m_mesh.LoadMesh(m_meshPath, (m_physicsParam.max + m_physicsParam.min) * 0.5, m_physicsParam.max - m_physicsParam.min, 1.0);
        m_meshMaterial.SetParam(kRadius * 4, 1.0);
        m_boundaryHash.SetCeilSize(kRadius * 4);
        m_boundaryHash.SetHashSize({64, 64, 64});
        Commit();
        m_boundaryHash.SetObjectSize(m_boundary.GetSize());
    }

    mps_physics physicsABI() const {
        const PhysicsParam& p = m_physicsParam;
        return mps_physics{p.dt, {p.gravity.x, p.gravity.y, p.gravity.z}, {p.min.x, p.min.y, p.min.z}, {p.max.x, p.max.y, p.max.z}};
    }

    mps_config m_cfg{};
    mps_ctx* m_ctx = nullptr;
    mps_step_stats m_stats{};
    bool b_runSim = false;
    bool m_physicsDirty = false;
    uint64_t m_frame = 0;
    uint32_t m_particleWSize = 5;
    std::string m_meshPath;
    PhysicsParam m_physicsParam;
    SPHMaterial m_sphMaterial;
    MeshMaterial m_meshMaterial;
    SPHObject m_sph;
    MeshObject m_mesh;
    BoundaryParticleObject m_boundary;
    SpatialHash m_sphHash, m_boundaryHash;
};

}  // namespace mps

#endif  // MPS_SYSTEM_HPP
