/*
 * sph_oracle.h — CPU restatement of the MPS_CUDA SPH step. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and only as the checker (or as the timed CPU baseline). The product path
 * (mps_cuda_amd) never links or calls it.
 *
 * The oracle mirrors the HIP library's C ABI (include/mps_sph.h) with an `orc_` prefix so
 * that parity tests are ABI-to-ABI. It is templated on the working precision:
 *   precision 64 — the reference's own arithmetic (REAL = double, src/MPS_Object/MPSDef.h:5-13);
 *   precision 32 — the FP32 semantics the HIP library implements: every REAL value and
 *                  literal is float, evaluated in the reference's expression order with no
 *                  FMA contraction; cell coordinates use trunc(double(x) * (1.0/double(cell)))
 *                  (MPSSpatialHash.cuh:9-19); convergence sums accumulate T-rounded terms in
 *                  double; solver scalars (CG alpha/beta, Jacobi omega) are computed in double
 *                  and rounded to T. Boundary sampling always runs in double and is stored as T.
 * Arrays exchanged through orc_get_array/orc_set_array are in the working precision
 * (float or double); colours are always float; ids/keys are uint32.
 *
 * Parity unpinned: the order of equal keys in the reference's thrust::sort_by_key
 * (MPSSpatialHash.cu:69,94,116,132; the Thrust of CUDA 12.6, not vendored with the reference).
 * It decides the ids within a cell and the Z-sorted payload order. This oracle defines it as
 * stable (ties by ascending prior index), checked against a brute-force stable sort; no
 * reference test or fixture holds the reference's own order.
 */
#ifndef MPS_SPH_ORACLE_H
#define MPS_SPH_ORACLE_H

#include "../include/mps_sph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

int orc_create(const mps_config* cfg, int precision, orc_ctx** out);
void orc_destroy(orc_ctx* c);
const char* orc_last_error(const orc_ctx* c);
int orc_precision(const orc_ctx* c);
int orc_set_threads(int n);

int orc_set_physics(orc_ctx* c, const mps_physics* p);
int orc_set_fluid_material(orc_ctx* c, const mps_sph_material* m);
/* Inputs in double; stored in the working precision. vel/mass may be NULL. */
int orc_upload_fluid(orc_ctx* c, size_t n, const double* pos, const double* vel,
                     const double* radius, const double* mass);
int orc_upload_boundary(orc_ctx* c, size_t n, const double* pos, const double* radius,
                        const double* mass);
int orc_sample_boundary_mesh(orc_ctx* c, const double* verts, size_t nv, const uint32_t* tris,
                             size_t nt, const double* center, const double* size,
                             const mps_mesh_material* m);
/* ParticleSampling again on the mesh of the last orc_sample_boundary_mesh with moved vertices
 * (same faces): faces whose sample count or shortest edge changed are generated anew. */
int orc_resample_boundary_mesh(orc_ctx* c, const double* verts, size_t nv, const mps_mesh_material* m,
                               size_t* regenerated);

/* The walls as k <= 4 boundary objects of the given sizes (contiguous, in order). */
int orc_set_boundary_objects(orc_ctx* c, uint32_t k, const uint32_t* counts);

/* A coupled world of fluid objects: nphase contexts (load each with the calls above through
 * orc_world_phase) whose density, factor, delta and pressure force include the other phases
 * (the reference's "ref" overloads); the pressure loops run jointly over all phases. */
typedef struct orc_world orc_world;
int orc_world_create(const mps_config* cfg, int precision, uint32_t nphase, orc_world** out);
void orc_world_destroy(orc_world* w);
orc_ctx* orc_world_phase(orc_world* w, uint32_t k);
int orc_world_step(orc_world* w, mps_step_stats* st /* [nphase] */);
int orc_world_run_phase(orc_world* w, int32_t phase, mps_step_stats* st /* [nphase] */);

int orc_step(orc_ctx* c, mps_step_stats* out);
int orc_run_phase(orc_ctx* c, int32_t phase, mps_step_stats* out);

int orc_array_info(orc_ctx* c, int32_t id, size_t* count, size_t* elem_bytes);
int orc_get_array(orc_ctx* c, int32_t id, void* dst, size_t bytes);
int orc_set_array(orc_ctx* c, int32_t id, const void* src, size_t bytes);
int orc_get_frame(orc_ctx* c, uint32_t* frame);
int orc_set_frame(orc_ctx* c, uint32_t frame);

/* Test hook: out = A(f) p with the implicit viscosity/surface-tension operator of the
 * current state (density and the pair lists must be current; p, out: N*3 in the working
 * precision). Runs the surface-tension factor and the pair-coefficient set-up first. */
int orc_cg_operator(orc_ctx* c, const void* p, void* out, double f);

/* Scalar helpers for known-answer tests (double in/out, evaluated in `precision`). */
int orc_kernel_eval(int precision, int kernel, size_t n, const double* dist, double invh,
                    double* w, double* g, double* stw, double* stg);
int orc_cell_keys(int precision, size_t n, const double* pos, double cell,
                  const uint32_t* hash_size, int morton, uint32_t* keys);
/* Mesh helpers: ownership bits (RTriangleBuild_kernel semantics) and per-face sample counts. */
int orc_mesh_rtri(size_t nv, const uint32_t* tris, size_t nt, uint32_t* rtri);
int orc_mesh_sample_counts(const double* verts, size_t nv, const uint32_t* tris, size_t nt,
                           double radius, uint32_t* counts);

#ifdef __cplusplus
}
#endif
#endif
