/*
 * mps_sph.h — C ABI of the MI355X (gfx950) SPH time step.
 *
 * This is the drop-in boundary for the compute that MPS_CUDA runs in MPS_Computer and
 * MPS_Object. The reference exposes that compute as mangled C++ DLL exports taking POD
 * bundles of raw device pointers (src/MPS_Computer/HeaderPre.h:1-19,
 * src/MPS_Computer/MPSSPHDFSPHUtil.h:9-87, src/MPS_Object/MPSSpatialHash.h:20-57).
 * Here every entry point is extern "C", takes plain pointers and sizes, returns an int
 * status (0 = OK) and never throws. A context owns all device memory and one HIP stream;
 * there is no global state, and a context is used by one host thread at a time.
 *
 * Reference entry points replaced (file:line of the reference interface):
 *   mps_create / mps_destroy         mps::System::System, SetDevice       MPSSystem.cpp:19,86-93
 *   mps_set_physics                  PhysicsParam (GBArchiver)             MPSUniformDef.h:19-25, MPSSystem.cpp:30-34
 *   mps_set_fluid_material           SPHMaterial::SetParam/SetRadius/...   MPSSPHMaterial.h:22-43, .cpp:4-44
 *   mps_upload_fluid                 SPHObject::Resize + CopyFromHost      MPSSystem.cpp:630-635
 *   mps_upload_boundary              BoundaryParticleObject::Resize        MPSBoundaryParticleObject.cu:14-24
 *   mps_sample_boundary_mesh         MeshObject::LoadMesh + kernel::ParticleSampling::ParticleSampling
 *                                    MPSMeshObject.cu:58-103, MPSParticleSamplingUtil.cu:42-133
 *   mps_set_boundary_objects         several BoundaryParticleObjects; boundary "ref" volume overload
 *                                    MPSSPHBaseUtil.cu:25-42
 *   mps_resample_boundary_mesh       ParticleSampling again on a moved mesh (IsSamplingParticleRequired,
 *                                    pIsGenerated / prevIdx)  MPSParticleSamplingUtil.cu:14-133
 *   mps_read_obj,                    the OBJ reading of MeshObject::LoadMesh  MPSMeshObject.cu:58-94
 *   mps_sample_boundary_obj          (+ the above for the file's mesh)
 *   mps_step                         System::OnUpdate                      MPSSystem.cpp:191-291
 *   mps_run_phase                    the kernel::SPH / kernel::Advect / SpatialHash calls made by
 *                                    OnUpdate, one at a time              MPSSystem.cpp:220-278
 *   mps_array_info/get/set           SpatialHash::GetHashKeys/GetHashIDs/GetStartIdx/GetEndIdx/
 *                                    GetNeighborhood and the SPHParam pointers
 *                                    MPSSpatialHash.h:33-38, MPSSPHParam.h:7-20
 *   mps_readback_async/fence_wait    CUDA-GL interop buffer map/unmap       MCUDAGLBuffer.inl:49-79,
 *                                    consumed by SPHRenderer::Draw         MPSSPHRenderer.cpp:18-32
 */
#ifndef MPS_SPH_H
#define MPS_SPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define MPS_OK 0
#define MPS_ERR_INVALID 1   /* bad argument or size */
#define MPS_ERR_HIP 2       /* a HIP runtime call failed */
#define MPS_ERR_NOMEM 3     /* device or host allocation failed */
#define MPS_ERR_STATE 4     /* call not valid in the current state (e.g. no fluid) */
#define MPS_ERR_NONFINITE 5 /* a NaN/inf was detected in the particle state */

/* Longest per-iteration error log kept in mps_step_stats (cap of 100 iterations + 1). */
#define MPS_MAX_ITER_LOG 128

/* ---- enums ------------------------------------------------------------------------ */
/* Smoothing kernel family. POLY6_SPIKY is the reference's active choice
 * (USE_CUBIC_KERNEL 0, MPSSPHKernel.cuh:5); CUBIC is the spline behind that switch
 * (MPSSPHKernel.cuh:34-40,52-59). */
enum mps_kernel_type { MPS_KERNEL_POLY6_SPIKY = 0, MPS_KERNEL_CUBIC = 1 };

/* Implicit viscosity / surface-tension solver. CG is the live reference path
 * (MPSSPHNonPressureUtil.cu:127-188); JACOBI is the 3x3 block-Jacobi viscosity solve that
 * exists only as commented-out code (SPHUtil.cu:344-453); NONE skips the phase. */
enum mps_visc_solver { MPS_VISC_CG = 0, MPS_VISC_JACOBI = 1, MPS_VISC_NONE = 2 };

/* Phases of one step, in OnUpdate order (MPSSystem.cpp:220-278). */
enum mps_phase {
    MPS_PHASE_ZSORT = 1,            /* SpatialHash::ZSort, boundary then fluid (:220-224) */
    MPS_PHASE_HASH = 2,             /* SpatialHash::UpdateHash, boundary then fluid (:226-227) */
    MPS_PHASE_NEIGHBORS = 3,        /* BuildNeighorhood bb, ff, fb (:229-231) */
    MPS_PHASE_BOUNDARY_VOLUME = 4,  /* ComputeBoundaryParticleVolumeSub/Final (:237-238) */
    MPS_PHASE_DENSITY = 5,          /* ResetREAL, ComputeDensitySub x2, Final, DensityColorTest (:240-245) */
    MPS_PHASE_ALPHA = 6,            /* ComputeDFSPHFactorSub x2, Final (:247-251) */
    MPS_PHASE_GRAVITY = 7,          /* ResetForce, ApplyGravity, UpdateVelocity (:253-255) */
    MPS_PHASE_DIVERGENCE = 8,       /* ComputeDFSPHDivergenceFree (:257-264) */
    MPS_PHASE_VISCOSITY = 9,        /* ApplyImplicitViscosityNSurfaceTension (:266) */
    MPS_PHASE_CONSTANT_DENSITY = 10,/* ComputeDFSPHConstantDensity (:268-275) */
    MPS_PHASE_ADVECT = 11           /* BoundaryCollision, UpdatePosition (:277-278) */
};

/* Arrays readable/writable for parity checks. Element layout in brackets; "real" is
 * float in the HIP library and the oracle's working precision in the oracle. */
enum mps_array_id {
    MPS_ARR_FLUID_POS = 1,       /* real[N][3]  SPHParam::pPosition */
    MPS_ARR_FLUID_VEL = 2,       /* real[N][3]  pVelocity */
    MPS_ARR_FLUID_RADIUS = 3,    /* real[N]     pRadius (particle radius r, support h = 4r) */
    MPS_ARR_FLUID_MASS = 4,      /* real[N]     pMass */
    MPS_ARR_FLUID_COLOR = 5,     /* float[N][4] pColor (DensityColorTest output) */
    MPS_ARR_FLUID_DENSITY = 6,   /* real[N]     pDensity */
    MPS_ARR_FLUID_ALPHA = 7,     /* real[N]     pFactorDFSPH */
    MPS_ARR_FLUID_KAPPA = 8,     /* real[N]     pPressure (stiffness of the last evaluation) */
    MPS_ARR_FLUID_DELTA = 9,     /* real[N]     pTempReal (density change rate of the last evaluation) */
    MPS_ARR_FLUID_FACTOR_ST = 10,/* real[N]     pFactorST */
    MPS_ARR_FLUID_CG_B = 11,     /* real[N][3]  pTempVec3 after ComputeSurfaceTensionCGb */
    MPS_ARR_BND_POS = 20,        /* real[B][3] */
    MPS_ARR_BND_RADIUS = 21,     /* real[B] */
    MPS_ARR_BND_VOLUME = 22,     /* real[B]     pVolume (Akinci volume) */
    MPS_ARR_BND_MASS = 23,       /* real[B] */
    MPS_ARR_BND_VEL = 24,        /* real[B][3] */
    MPS_ARR_BND_FACE_ID = 25,    /* uint32[B] */
    MPS_ARR_BND_BCC = 26,        /* real[B][2]  barycentric weights (w0, w1) on the owning face */
    MPS_ARR_BND_OBJECT = 27,     /* uint32[B]   wall object of each wall particle (mps_set_boundary_objects) */
    MPS_ARR_FHASH_KEY = 30,      /* uint32[N]   sorted cell keys of the fluid hash */
    MPS_ARR_FHASH_ID = 31,       /* uint32[N]   particle ids in key order */
    MPS_ARR_FHASH_START = 32,    /* uint32[C]   first sorted slot of each cell, 0xFFFFFFFF if empty */
    MPS_ARR_FHASH_END = 33,      /* uint32[C]   one past the last slot, 0xFFFFFFFF if empty */
    MPS_ARR_BHASH_KEY = 34,
    MPS_ARR_BHASH_ID = 35,
    MPS_ARR_BHASH_START = 36,
    MPS_ARR_BHASH_END = 37,
    MPS_ARR_NEI_FF_IDX = 40,     /* uint32[N+1] CSR offsets, fluid->fluid (self excluded) */
    MPS_ARR_NEI_FF = 41,         /* uint32[nnz] neighbour ids in 27-cell z,y,x order */
    MPS_ARR_NEI_FB_IDX = 42,     /* uint32[N+1] fluid->boundary */
    MPS_ARR_NEI_FB = 43,
    MPS_ARR_NEI_BB_IDX = 44,     /* uint32[B+1] boundary->boundary (self excluded) */
    MPS_ARR_NEI_BB = 45
};

/* ---- parameter structs -------------------------------------------------------------- */
/* PhysicsParam (MPSUniformDef.h:19-25): time step, gravity and collision box. */
typedef struct mps_physics {
    double dt;
    double gravity[3];
    double box_min[3];
    double box_max[3];
} mps_physics;

/* SPHMaterialParam (MPSSPHMaterial.h:9-21). `radius` is the material radius R, equal to
 * the support h = 4 r of the particles; volume = R^3/48*3.141592 and mass = density*volume
 * are derived exactly as SPHMaterial::SetParam does (MPSSPHMaterial.cpp:13-19). */
typedef struct mps_sph_material {
    double radius;
    double density;
    double viscosity;
    double xsph;
    double surface_tension;
} mps_sph_material;

/* MeshMaterialParam (MPSMeshMaterial.h:9-20): sampling radius R_m; boundary particles get
 * radius R_m/4 and spacing R_m/2 (MPSParticleSamplingUtil.cuh:32,92). */
typedef struct mps_mesh_material {
    double radius;
    double density;
} mps_mesh_material;

/* Everything the reference hard-codes as defines or literals, in one struct.
 * mps_config_init fills the reference's own values. */
typedef struct mps_config {
    int32_t kernel;              /* enum mps_kernel_type; reference: POLY6_SPIKY */
    int32_t visc_solver;         /* enum mps_visc_solver; reference: CG */
    int32_t faithful_quirks;     /* 1: boundary volume accumulates across steps and is not
                                    permuted by the Z-sort (MPSSystem.cpp:237-238,
                                    MPSSpatialHash.cu:132-139); 0: V_b = 1/S_b each step */
    int32_t device;              /* HIP device ordinal (HIP library only) */
    uint32_t hash_size[3];       /* cells per axis, powers of two; reference 64^3 */
    double cell_size;            /* hash cell edge; <=0 means "fluid material radius" (= h) */
    double df_tol;               /* divergence-free error tolerance; reference 1e-3 */
    uint32_t df_min_iter;        /* break only when l > df_min_iter; reference 1 */
    double cd_tol;               /* constant-density tolerance; reference 1e-4 */
    uint32_t cd_min_iter;        /* reference 2 */
    uint32_t max_iter;           /* DF/CD cap; reference 100 */
    double cg_tol;               /* absolute r.r threshold (reference 1e-2); < 0: -cg_tol * r0.r0 */
    uint32_t cg_max_iter;        /* reference 100 */
    double cg_factor;            /* diagonal shift f used after the first A.x; reference 0.01 */
    uint32_t jacobi_max_iter;    /* Jacobi cap; reference 100 */
    double jacobi_tol;           /* Jacobi stop on mean error*dt/N; reference 1e-4 */
    uint32_t zsort_interval;     /* Z-sort when frame % interval == 1; reference 50; 0 = never */
    uint32_t nei_capacity;       /* initial per-particle neighbour slots (grows on overflow) */
    uint32_t host_sync_every;    /* HIP library: solver iterations enqueued between
                                    convergence checks on the host (0 = adaptive) */
    uint32_t adj_cg_max_iter;    /* adjoint: cap of the transposed viscosity solve (300) */
    uint32_t use_graphs;         /* HIP library: replay solver iteration chunks from cached HIP
                                    graphs (same results; not while timing kernels) */
    uint32_t halo_limit;         /* HIP library: largest block halo listed for LDS staging; blocks
                                    with more gather from global memory in every halo-staged kernel.
                                    0 = the largest listed (1536; each kernel then stages up to its
                                    own LDS depth); smaller values exist to test the global path */
    uint32_t couple_scenes;      /* 1: the scenes of a batch are fluid objects of one world (same dt,
                                    gravity, box; walls given per scene): density, factor, density
                                    change and pressure force include the other scenes' particles
                                    (the reference's "ref" overloads, MPSSPHBaseUtil.cu:73-92,
                                    MPSSPHDFSPHUtil.cu:31-51,105-127,212-234) and the pressure loops
                                    run jointly, ending when every scene meets its tolerance;
                                    viscosity and surface tension stay within each scene. 0: the
                                    scenes are independent. */
    uint32_t cg_stream_coeffs;   /* HIP library: 1 = the viscosity operator streams both pair
                                    coefficients (a_ij, b_ij) per iteration (8 bytes per pair);
                                    0 = it streams a_ij and recomputes b_ij from the positions
                                    (4 bytes per pair, one square root and one division). Same
                                    results; mps_autotune may change it. */
    uint32_t persistent_solver;  /* HIP library: nonzero = the viscosity CG runs all its iterations in
                                    one cooperative launch (grid-wide barriers, every workgroup folds
                                    the sums itself) when every workgroup of the context fits the
                                    device at once; otherwise, and under kernel timing, the chunked
                                    launches. 1: barriers with agent-scope fences (each workgroup
                                    writes back and invalidates its L2); 2: the vectors and sums that
                                    cross workgroups are read and written device-coherently, and the
                                    barrier only counts arrivals. Same results; mps_autotune may
                                    change it (it tries 0 and 2). */
    uint32_t persistent_spin;    /* polls before a barrier wait of the persistent solver gives up
                                    (the solve then reruns with chunked launches); 0 = 4194304 */
    double adj_cg_tol;           /* adjoint: its r.r threshold, same convention as cg_tol (-1e-12) */
    uint32_t cg_fold_in_kernel;  /* HIP library: 1 = the chunked viscosity CG folds its sums in the last
                                    workgroup of the A.p and update launches (two launches per
                                    iteration; arrivals counted with device-scope atomics); 0 =
                                    separate one-workgroup fold launches (four; the default). Same
                                    results; mps_autotune may change it. The allocation-free step
                                    always uses the fold launches. */
    uint32_t cg_halo_depth;      /* HIP library: LDS depth of the streamed viscosity operator's block
                                    halos. 0 = chosen each step from the previous list build (the
                                    deep form once more than 1 in 1000 workgroups have halos over
                                    the shallow depth); 1 = shallow (1264 records, five workgroups
                                    per CU); 2 = deep (1536, four per CU). A deeper halo is gathered
                                    from global memory. Same results. */
} mps_config;

/* Per-step statistics (the reference prints these to the debugger only:
 * MPSSPHDFSPHUtil.cu:335-339,378-383,421-425,435-440, MPSSolver.cuh:44-48). */
typedef struct mps_step_stats {
    uint32_t frame;        /* frame counter before the step (System::m_frame) */
    uint32_t zsorted;      /* 1 if the Z-sort ran in this step */
    uint32_t l_df;         /* value of the loop variable l when the DF loop ended */
    uint32_t l_cd;
    uint32_t l_cg;         /* CG loop count (MPSSolver.cuh return value) or Jacobi sweeps */
    uint32_t hit_cap_df;   /* 1 if the loop ran to its cap without converging */
    uint32_t hit_cap_cd;
    uint32_t hit_cap_cg;
    uint32_t nonfinite;    /* 1 if a NaN/inf was found in positions/velocities after the step */
    uint32_t n_fluid;
    uint32_t n_boundary;
    uint32_t nei_regrow;   /* number of neighbour-capacity grow-and-retry events */
    uint64_t nnz_ff;
    uint64_t nnz_fb;
    uint64_t nnz_bb;
    double err_df[MPS_MAX_ITER_LOG];  /* mean positive error of each DF evaluation */
    double err_cd[MPS_MAX_ITER_LOG];
    double rr_cg[MPS_MAX_ITER_LOG];   /* r.r of each CG iteration (or Jacobi error) */
    /* block halos of the fluid lists (no reference counterpart): largest halo listed for LDS
     * staging (at most 1536), 256-slot blocks over that limit (these gather from global memory),
     * blocks with particles; zero when the CG did not run */
    uint32_t halo_max;
    uint32_t halo_spill;
    uint32_t halo_blocks;
    /* how the viscosity CG ran (HIP library): 0 chunked launches, 1 one persistent cooperative
     * launch, 2 the persistent launch gave up at a barrier and the solve reran chunked */
    uint32_t cg_persistent;
} mps_step_stats;

typedef struct mps_ctx mps_ctx;

/* Completion handle of an asynchronous readback. */
typedef struct mps_fence {
    void* event;       /* opaque */
    uint64_t frame;
} mps_fence;

/* ---- lifetime --------------------------------------------------------------------- */
int mps_config_init(mps_config* cfg);
int mps_create(const mps_config* cfg, mps_ctx** out);
void mps_destroy(mps_ctx* ctx);
const char* mps_last_error(const mps_ctx* ctx);
const char* mps_version(void);

/* ---- scene setup ------------------------------------------------------------------- */
int mps_set_physics(mps_ctx* ctx, const mps_physics* p);
int mps_set_fluid_material(mps_ctx* ctx, const mps_sph_material* m);
/* pos, vel: n*3 floats; radius, mass: n floats. vel may be NULL (zero); mass may be NULL
 * (material mass, MPSSystem.cpp:583). Resets the frame counter and all derived state. */
int mps_upload_fluid(mps_ctx* ctx, size_t n, const float* pos, const float* vel,
                     const float* radius, const float* mass);
/* Pre-sampled boundary particles: pos n*3, radius n. mass may be NULL (zero). */
int mps_upload_boundary(mps_ctx* ctx, size_t n, const float* pos, const float* radius,
                        const float* mass);
/* Triangle mesh -> boundary particles. verts: nv*3 doubles, tris: nt*3 vertex indices.
 * If center/size are non-NULL the mesh is first fitted into that box as
 * MeshObject::Translate does (MPSMeshObject.cu:130-151). The vertex mass is
 * total area * density / nv (MPSMeshObject.cu:169-174). */
int mps_sample_boundary_mesh(mps_ctx* ctx, const double* verts, size_t nv,
                             const uint32_t* tris, size_t nt, const double* center,
                             const double* size, const mps_mesh_material* m);
/* The wall particles as k <= 4 boundary objects (several obstacles, each its own
 * BoundaryParticleObject in the reference): counts[o] particles each, contiguous, in the
 * current order (e.g. upload the objects' walls concatenated). A wall's volume then sums its
 * own object first and every other object after it, in object order (the boundary "ref"
 * overload, ComputeBoundaryParticleVolumeSub(boundary, nei, refBoundary),
 * MPSSPHBaseUtil.cu:25-42); the fluid's wall sums in density, factor, density change and
 * pressure force are taken one object after the other; each object is Z-sorted within its own
 * range. Call after the walls are uploaded or sampled; replacing the walls makes them one
 * object again. Single-scene contexts only; the adjoint refuses several objects. */
int mps_set_boundary_objects(mps_ctx* ctx, uint32_t k, const uint32_t* counts);
/* ParticleSampling called again on the mesh of the last mps_sample_boundary_mesh after its
 * vertices moved (MPSParticleSamplingUtil.cu:14-133, .cuh:69-246). `verts` are the nv new
 * vertex positions (same count and faces; no fitting), m->radius the sampling radius. The
 * sample count and shortest edge of every face are kept from the previous sampling; a face
 * where either changed is generated anew, every other face keeps its previous particles (moved
 * to the face's new offset, positions unchanged). Nothing changes when no face changed. As in
 * the reference, the particle arrays are resized by index: with faithful_quirks a kept
 * particle takes the mass, velocity and volume stored at its new index (new slots start at
 * zero); without, it carries its own mass and velocity. The vertex masses stay those of the
 * first sampling. *regenerated (may be NULL) receives the number of faces generated anew.
 * MPS_ERR_STATE if no mesh was sampled or the walls were replaced since. */
int mps_resample_boundary_mesh(mps_ctx* ctx, const double* verts, size_t nv,
                               const mps_mesh_material* m, size_t* regenerated);
/* The file-reading half of MeshObject::LoadMesh (MPSMeshObject.cu:58-103): the vertices
 * (x, y, z) and 0-based triangles of a Wavefront OBJ file, read token by token as the
 * reference does (tokens longer than one character are skipped; `v` takes three reals, `f`
 * three 1-based indices; a quad keeps its first triangle; `a/b/c` indices use `a`). Call
 * with verts = tris = NULL for the counts, then with buffers of at least *nv x 3 doubles and
 * *nt x 3 indices. Malformed numbers, indices of 0 or past the vertex count, an unreadable
 * file or one without vertices or faces give MPS_ERR_INVALID (the reference loops or exits).
 * `ctx` may be NULL (no GPU needed; no error message is kept then). */
int mps_read_obj(mps_ctx* ctx, const char* path, double* verts, size_t* nv, uint32_t* tris,
                 size_t* nt);
/* MeshObject::LoadMesh + ParticleSampling from an OBJ file: mps_read_obj, then
 * mps_sample_boundary_mesh with the same center/size/material arguments. */
int mps_sample_boundary_obj(mps_ctx* ctx, const char* path, const double* center,
                            const double* size, const mps_mesh_material* m);

/* ---- stepping ---------------------------------------------------------------------- */
/* One System::OnUpdate. `out` may be NULL. */
int mps_step(mps_ctx* ctx, mps_step_stats* out);
/* One mps_step that allocates no device memory (no counterpart in the reference; for callers
 * that cannot allocate: tight serving loops, harnesses that forbid device allocation). The
 * step's reduction scratch (per-workgroup FP64 partial sums, neighbour-count records, solver
 * records, loop counters) lives in a library-wide device array of 1.5 MiB (held on each GPU the
 * library runs on until the process ends), sized for at most 8,388,608 fluid slots or wall
 * particles and 64 scenes (MPS_ERR_INVALID beyond). Every other buffer is the context's and must
 * already be large enough: after one mps_step on a scene at least this large (and one Z-sort
 * step, if this step Z-sorts); otherwise, or when the neighbour lists outgrow their capacity,
 * the call fails with MPS_ERR_STATE before any physics is applied, and an mps_step continues
 * from the unchanged state. Solver chunks are launched directly (no graph capture) and kernel
 * timing must be off. Calls on one GPU share its scratch array: calls that overlap on one GPU
 * (several threads, several contexts) race. Results equal mps_step's bit for bit. */
int mps_step_noalloc(mps_ctx* ctx, mps_step_stats* out);
/* Run one phase in isolation on the current state (teacher-forced parity checks).
 * Solver phases fill the matching fields of `out` (may be NULL). */
int mps_run_phase(mps_ctx* ctx, int32_t phase, mps_step_stats* out);
int mps_synchronize(mps_ctx* ctx);

/* ---- state access -------------------------------------------------------------------- */
int mps_array_info(mps_ctx* ctx, int32_t id, size_t* count, size_t* elem_bytes);
/* Blocking copies of a whole array; `bytes` must equal count*elem_bytes. */
int mps_get_array(mps_ctx* ctx, int32_t id, void* dst, size_t bytes);
int mps_set_array(mps_ctx* ctx, int32_t id, const void* src, size_t bytes);
int mps_get_frame(mps_ctx* ctx, uint32_t* frame);
int mps_set_frame(mps_ctx* ctx, uint32_t frame);

/* ---- render feed (replaces the CUDA-GL interop buffers) -------------------------------- */
/* Device->host copies of positions (n*3 float), radii (n) and colours (n*4) in reference
 * order; any pointer may be NULL. The state is packed on the context's stream into one of two
 * device staging slots (alternating per call) and copied on a second stream of the context
 * behind an event, so the next mps_step's kernels run while the copy drains (step N+1 overlaps
 * the copy of frame N). Host buffers must be pinned (mps_host_alloc) for the copy to be
 * asynchronous, and must not be reused until mps_fence_wait on the fence returns; use two sets
 * of host buffers to keep two frames in flight. Replaces mapping the GL buffers to CUDA
 * (MCUDAGLBuffer.inl:49-79); a caller uploads the host buffers to its renderer. */
int mps_readback_async(mps_ctx* ctx, float* host_pos, float* host_radius, float* host_rgba,
                       mps_fence* fence);
/* The same in the layout SPHRenderer's shader reads (src/shader/particleShader.vert:3-5):
 * positions (n*3) and radii (n) as double, colours (n*4) as float. */
int mps_readback_async_f64(mps_ctx* ctx, double* host_pos, double* host_radius, float* host_rgba,
                           mps_fence* fence);
/* Wait until the fence's copies are in the host buffers. A fence stays valid after later
 * readbacks (waiting then also covers any later copy of the same staging slot). */
int mps_fence_wait(mps_ctx* ctx, mps_fence* fence);
int mps_host_alloc(mps_ctx* ctx, size_t bytes, void** out);
int mps_host_free(mps_ctx* ctx, void* p);

/* ---- batched scenes ---------------------------------------------------------------------
 * No counterpart in the reference, which steps one scene per System. A context can hold S
 * independent scenes (own particles, walls, physics and material) stored back to back; every
 * phase of mps_step is then one launch over all of them, each solver loop keeps a record and a
 * convergence flag per scene, and a scene whose loop has converged sits out the remaining
 * iterations. Each scene's results are those of the same scene stepped alone in a context of
 * its own with the same configuration. mps_upload_* and mps_sample_boundary_mesh are
 * single-scene calls and fail with MPS_ERR_STATE while a batch is held; mps_set_physics and
 * mps_set_fluid_material then apply to every scene. Fluid arrays are exported scene after
 * scene (mps_scene_info gives each scene's range); hash tables only for one scene. */
typedef struct mps_scene_range {
    size_t fluid_offset;    /* reference index of the scene's first fluid particle */
    size_t n_fluid;
    size_t boundary_offset; /* index of its first boundary particle */
    size_t n_boundary;
} mps_scene_range;

/* Drop every scene and all particles. */
int mps_batch_clear(mps_ctx* ctx);
/* Stage one scene (copied; uploaded by mps_batch_commit). pos/vel n_fluid*3, radius/mass
 * n_fluid (vel NULL = 0, mass NULL = material mass); boundary bpos n_boundary*3, bradius and
 * bmass n_boundary (bmass NULL = 0). n_fluid must be > 0. */
int mps_batch_add_scene(mps_ctx* ctx, const mps_physics* phys, const mps_sph_material* mat,
                        size_t n_fluid, const float* pos, const float* vel, const float* radius,
                        const float* mass, size_t n_boundary, const float* bpos,
                        const float* bradius, const float* bmass);
/* Upload the staged scenes in the order added; replaces the context's state, frame = 0. */
int mps_batch_commit(mps_ctx* ctx);
int mps_scene_count(mps_ctx* ctx, uint32_t* n);
int mps_scene_info(mps_ctx* ctx, uint32_t scene, mps_scene_range* out);
/* Statistics of one scene in the last step (iteration counts, error logs, caps). */
int mps_scene_stats(mps_ctx* ctx, uint32_t scene, mps_step_stats* out);
/* Change one scene's material (viscosity, surface tension, ...) between steps. */
int mps_scene_set_material(mps_ctx* ctx, uint32_t scene, const mps_sph_material* m);

/* ---- adjoint ------------------------------------------------------------------------------
 * No counterpart in the reference. Reverse mode through mps_step with each step's control
 * flow frozen: its neighbour lists, the DF/CD iteration counts (differentiated through their
 * iterations), the per-particle branches; the viscosity / surface-tension solve differentiated
 * implicitly (transposed solve with the same operator). Either kernel set (cubic spline, or
 * poly6/spiky) and any fluid radii; needs the CG viscosity solve (MPS_ERR_STATE otherwise, and for
 * coupled scenes or several wall objects). A relative CG tolerance (cg_tol < 0) makes the forward
 * solve match the implicit derivative. */
/* Step n_steps times from the current state, keeping a checkpoint every checkpoint_every
 * steps. loop_counts (may be NULL): n_steps * scenes * 3 values (l_df, l_cg, l_cd per scene). */
int mps_adj_rollout(mps_ctx* ctx, uint32_t n_steps, uint32_t checkpoint_every, uint32_t* loop_counts);
/* mps_adj_rollout that also returns the states after steps obs_step[0..n_obs) (each in
 * 1..n_steps): positions into pos_out and velocities into vel_out (may be NULL), n_obs blocks of
 * n*3 floats in the reference order of each state. */
int mps_adj_rollout_traj(mps_ctx* ctx, uint32_t n_steps, uint32_t checkpoint_every, uint32_t n_obs,
                         const uint32_t* obs_step, float* pos_out, float* vel_out, uint32_t* loop_counts);
/* Gradients of a loss L(final positions, final velocities) given dL/dpos (n*3, reference
 * order) and dL/dvel (n*3 or NULL): dL/d(velocities at the rollout's start) into dloss_dvel0
 * (n*3), and per scene dL/dviscosity and dL/dsurface_tension (may be NULL). Steps are
 * recomputed from the checkpoints; the context is left at the rollout's final state. */
int mps_adj_backward(mps_ctx* ctx, const float* dloss_dpos, const float* dloss_dvel, float* dloss_dvel0,
                     double* dloss_dviscosity, double* dloss_dsurface_tension);
/* The same for a loss of the whole trajectory: the states after steps obs_step[0..n_obs) of the
 * rollout (each in 1..n_steps; any order; a repeated step adds) with cotangents dloss_dpos
 * (n_obs blocks of n*3, observation k at offset k*n*3, each in the reference order of its
 * observed state) and dloss_dvel (same layout, or NULL). mps_adj_backward is the case n_obs = 1,
 * obs_step = {n_steps}. MPS_ERR_STATE when there is no rollout, or when particles, walls,
 * scenes, physics or materials were replaced since it. */
int mps_adj_backward_traj(mps_ctx* ctx, uint32_t n_obs, const uint32_t* obs_step, const float* dloss_dpos,
                          const float* dloss_dvel, float* dloss_dvel0, double* dloss_dviscosity,
                          double* dloss_dsurface_tension);

/* ---- autotuning ----------------------------------------------------------------------------
 * Times `steps` steps of the current state under each launch setting that leaves results
 * unchanged and keeps the fastest in the context's configuration, one dimension after the
 * other: graph replay of solver chunks on/off with host polling every 1, 4 or 16 iterations or
 * adaptive; then the viscosity operator's coefficient stream (both coefficients, or a_ij with
 * b_ij recomputed: 24- or 32-byte LDS records); then the CG's folds (separate launches, or in
 * the last workgroup of each launch); then the persistent CG solver where the grid fits the
 * device. The state is restored after every candidate. */
typedef struct mps_tune_result {
    uint32_t use_graphs;        /* chosen */
    uint32_t host_sync_every;   /* chosen (0 = adaptive) */
    uint32_t n_tried;
    uint32_t pad;
    double ms_per_step[8];      /* per candidate: graphs off with polling {0,1,4,16}, then on */
    uint32_t cg_stream_coeffs;  /* chosen with the launch setting above */
    uint32_t persistent_solver; /* chosen last (1 only where the grid fits the device) */
    double ms_coeffs[2];        /* cg_stream_coeffs 1, 0 */
    double ms_persistent[2];    /* persistent_solver 0, 2 (2: 0.0 when not applicable) */
    uint32_t cg_fold_in_kernel; /* chosen before the persistent solver */
    uint32_t pad2;
    double ms_fold[2];          /* cg_fold_in_kernel 0, 1 */
} mps_tune_result;
int mps_autotune(mps_ctx* ctx, uint32_t steps, mps_tune_result* out);

/* ---- per-kernel timing (replaces MTimer's debugger prints, MTimer.h:12-80) ---------------
 * When enabled, every kernel launch is bracketed by HIP events on the context's stream.
 * `launches` counts every launch; `useful` counts launches that did work (solver kernels
 * enqueued past convergence exit at once and are not useful). Enabling resets the counters. */
typedef struct mps_kernel_stat {
    char name[48];
    uint64_t launches;
    uint64_t useful;
    double total_ms;
} mps_kernel_stat;
int mps_profile_enable(mps_ctx* ctx, int32_t on);
/* Copies of `bytes` from one device buffer to another on the context's stream, `reps` times
 * each: first k_calib_copy (a plain grid-stride copy, the reference for scaling hardware byte
 * counters), then k_stream_copy (one 16-byte vector per thread, non-temporal: the copy rate of
 * the memory system). With kernel timing on, both appear in mps_profile_read. */
int mps_calibration_copy(mps_ctx* ctx, size_t bytes, int32_t reps);
int mps_profile_read(mps_ctx* ctx, mps_kernel_stat* out, size_t capacity, size_t* count);

#ifdef __cplusplus
}
#endif
#endif /* MPS_SPH_H */
