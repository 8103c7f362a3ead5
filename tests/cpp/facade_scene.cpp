// Drives libmps_sph.so through the C++ classes of include/mps_system.hpp the way the
// reference's application drives its System (MPSSystem.cpp:19-291): Initalize, the simulation
// keys, OnUpdate per frame.
//
//   facade_scene host <mesh.obj>
//       no GPU: prints the SPHMaterial / MeshMaterial defaults and derived values and the
//       mesh LoadMesh read, one "name value..." line each
//   facade_scene lists <mesh.obj> <out.bin>
//       Initalize, then the SpatialHash calls (UpdateHash, BuildNeighorhood, GetNeighborhood
//       for the fluid, fluid->wall and wall lists); writes them (format at lists())
//   facade_scene run <mesh.obj> <keys> <steps> <out.bin>
//       Initalize, then OnKeyDown for each key character ('+' = ']', '-' = '[', 'g' = 'G'),
//       ' ' to start, <steps> OnUpdate calls; with 'w' among the keys the +x wall of the box
//       moves out by 1 m after half of the steps (MeshObject::SetVertexPositions, then
//       ParticleSampling again); writes <out.bin>:
//       uint64 n, uint64 frame, uint64 n_boundary, uint32 l_df, l_cd, l_cg, pad,
//       float pos[n][3]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "mps_system.hpp"

static int host_checks(const char* mesh) {
    mps::SPHMaterial sm;
    std::printf("sph_default %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", sm.GetRadius(), sm.GetDensity(), sm.GetVolume(),
                sm.GetMass(), sm.GetViscosity(), sm.GetXSPH(), sm.GetSurfaceTension());
    sm.SetRadius(0.4);
    std::printf("sph_radius_0.4 %.17g %.17g %.17g\n", sm.GetRadius(), sm.GetVolume(), sm.GetMass());
    sm.SetParam(0.04, 1000.0);
    std::printf("sph_param_0.04_1000 %.17g %.17g %.17g\n", sm.GetRadius(), sm.GetVolume(), sm.GetMass());
    mps::MeshMaterial mm;
    std::printf("mesh_default %.17g %.17g %.17g\n", mm.GetParam().radius, mm.GetParam().density, mm.GetParam().mass);
    mps::MeshObject mo;
    mo.LoadMesh(mesh, mps::REAL3{1.0, 2.0, 3.0}, mps::REAL3{4.0, 5.0, 6.0}, 2.5);
    std::printf("mesh_counts %zu %zu\n", mo.GetVertexSize(), mo.GetSize());
    std::printf("mesh_verts");
    for (double v : mo.GetVertices()) std::printf(" %.17g", v);
    std::printf("\nmesh_faces");
    for (uint32_t f : mo.GetFaces()) std::printf(" %u", f);
    std::printf("\nmesh_fit %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", mo.GetCenter().x, mo.GetCenter().y, mo.GetCenter().z,
                mo.GetFitSize().x, mo.GetFitSize().y, mo.GetFitSize().z, mo.GetDensity());
    try {
        mps::MeshObject bad;
        bad.LoadMesh(std::string(mesh) + ".missing", mps::REAL3{}, mps::REAL3{}, 1.0);
        std::printf("missing_mesh no_error\n");
    } catch (const mps::Exception& e) {
        std::printf("missing_mesh status %d\n", e.code());
    }
    return 0;
}

static int run(const char* mesh, const char* keys, int steps, const char* out) {
    mps::System sys;
    sys.SetDevice(0);
    sys.Initalize(mesh);
    for (const char* k = keys; *k; k++) {
        if (*k == '+') sys.OnKeyDown(221);
        else if (*k == '-') sys.OnKeyDown(219);
        else if (*k == 'g') sys.OnKeyDown('G');
    }
    sys.OnUpdate();  // paused: no step
    if (sys.GetFrame() != 0) {
        std::fprintf(stderr, "OnUpdate stepped while paused\n");
        return 1;
    }
    sys.OnKeyDown(32);
    const bool move_wall = std::strchr(keys, 'w') != nullptr;
    for (int s = 0; s < steps; s++) {
        if (move_wall && s == steps / 2) {
            // the mesh is the unit cube fitted into the collision box: x = (v - 0.5) * size + center
            mps::MeshObject& walls = sys.GetMeshObject();
            const mps::REAL3 c = walls.GetCenter(), sz = walls.GetFitSize();
            std::vector<double> w(walls.GetVertices());
            for (size_t i = 0; i < w.size() / 3; i++) {
                const bool xmax = w[3 * i] == 1.0;
                w[3 * i] = (w[3 * i] - 0.5) * sz.x + c.x + (xmax ? 1.0 : 0.0);
                w[3 * i + 1] = (w[3 * i + 1] - 0.5) * sz.y + c.y;
                w[3 * i + 2] = (w[3 * i + 2] - 0.5) * sz.z + c.z;
            }
            walls.SetVertexPositions(w);
            mps::kernel::ParticleSampling::ParticleSampling(sys.GetMeshMaterial().GetParam(), walls,
                                                            sys.GetBoundaryParticleObject());
        }
        sys.OnUpdate();
    }
    const std::vector<float> pos = sys.GetPositions();
    const uint64_t hdr[3] = {pos.size() / 3, sys.GetFrame(), sys.GetBoundaryParticleObject().GetSize()};
    const mps_step_stats& st = sys.GetLastStats();
    const uint32_t loops[4] = {st.l_df, st.l_cd, st.l_cg, 0u};
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    std::fwrite(hdr, sizeof(hdr), 1, f);
    std::fwrite(loops, sizeof(loops), 1, f);
    std::fwrite(pos.data(), sizeof(float), pos.size(), f);
    std::fclose(f);
    std::printf("frames %llu particles %llu walls %llu\n", (unsigned long long)hdr[1], (unsigned long long)hdr[0],
                (unsigned long long)hdr[2]);
    return 0;
}

// SpatialHash::UpdateHash / BuildNeighorhood / GetNeighborhood and the hash getters on the
// Initalize scene: writes uint64 counts (keys, ff offsets, ff ids, fb offsets, fb ids, bb
// offsets, bb ids), then the uint32 arrays in that order.
static int lists(const char* mesh, const char* out) {
    mps::System sys;
    sys.SetDevice(0);
    sys.Initalize(mesh);
    mps::SpatialHash& h = sys.GetSPHHash();
    h.UpdateHash();
    h.BuildNeighorhood();
    const std::vector<uint32_t> ids = h.GetHashIDs();
    const auto ff = h.GetNeighborhood();
    const auto fb = h.GetNeighborhood(true);
    const auto bb = sys.GetBoundaryParticleHash().GetNeighborhood();
    const std::vector<uint32_t>* arrs[7] = {&ids, &ff.first, &ff.second, &fb.first, &fb.second, &bb.first, &bb.second};
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    for (const auto* a : arrs) {
        const uint64_t n = a->size();
        std::fwrite(&n, sizeof(n), 1, f);
    }
    for (const auto* a : arrs) std::fwrite(a->data(), sizeof(uint32_t), a->size(), f);
    std::fclose(f);
    std::printf("ids %zu ff %zu fb %zu bb %zu\n", ids.size(), ff.second.size(), fb.second.size(), bb.second.size());
    return 0;
}

int main(int argc, char** argv) {
    try {
        if (argc == 3 && !std::strcmp(argv[1], "host")) return host_checks(argv[2]);
        if (argc == 4 && !std::strcmp(argv[1], "lists")) return lists(argv[2], argv[3]);
        if (argc == 6 && !std::strcmp(argv[1], "run")) return run(argv[2], argv[3], std::atoi(argv[4]), argv[5]);
    } catch (const mps::Exception& e) {
        std::fprintf(stderr, "mps::Exception %d: %s\n", e.code(), e.what());
        return 2;
    }
    std::fprintf(stderr, "usage: facade_scene host <mesh.obj> | lists <mesh.obj> <out.bin> | run <mesh.obj> <keys> <steps> <out.bin>\n");
    return 64;
}
