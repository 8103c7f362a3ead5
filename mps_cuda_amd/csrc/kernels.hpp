// kernels.hpp — host-side launchers for the gfx950 kernels and the small POD structs
// they share with the context. No torch types, no CUDA types: plain HIP.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <tuple>
#include <type_traits>
#include <utility>

#include "../../include/mps_sph.h"

namespace mps {
namespace dev {
struct HashGeom;
struct f3;
}

// Read-only view of one set's hash (cell table + ids and positions in key order). For the
// fluid, whose working arrays are kept in key order, `id` is null and candidates are slots.
struct HashView {
    const uint32_t* start;
    const uint32_t* end;
    const uint32_t* id;
    const float4* spos;
    uint32_t n;
};

// Neighbour statistics written by the search kernels.
struct NbrCounters {
    uint32_t max_ff, max_fb, max_bb, max_fr;
    unsigned long long nnz_ff, nnz_fb, nnz_bb, nnz_fr;
};

// Barrier words of the persistent solver kernels (zeroed before each launch): arrivals per group
// of kGridGroup workgroups, arrivals of whole groups, the barrier generation, the give-up flag.
constexpr uint32_t kGridGroup = 16;
constexpr uint32_t kGridGroups = 128;  // up to 2048 workgroups
struct GridSync {
    uint32_t group[kGridGroups];
    uint32_t count, gen, gave_up, pad;
};
// Device-resident state of one iterative solve. The fold kernels finalise the sums here,
// so the host only needs to look at `done`/`l`.
struct SolverDev {
    uint32_t done;     // loop finished (converged); later kernels of the chunk exit at once
    uint32_t l;        // reference loop variable when the loop ended (or next iteration)
    uint32_t counter;  // workgroups of the current launch that arrived (folds in the last workgroup)
    uint32_t pad;
    double rr;         // CG: r.r of the current residual
    double pAp;        // CG: p.Ap
    double thr;        // CG: stop when r.r < thr (tol, or -tol * initial r.r when tol < 0)
    float alpha, beta; // CG step scalars (rounded to float like the FP32 reference)
    double log[MPS_MAX_ITER_LOG];
};

// Parameters of one scene. A context holds S >= 1 scenes stored back to back: scene s owns
// the fluid slots [begin, end), and `begin` is a multiple of the workgroup size, so every
// workgroup of a fluid kernel belongs to exactly one scene (its partial sums, and their fold,
// are then the same as when the scene runs alone).
struct StepParams {
    uint32_t N;           // fluid slots of the context (the column stride of the ELL lists)
    uint32_t begin, end;  // this scene's slots
    uint32_t count;       // its particles (end - begin)
    uint32_t blk0, nblk;  // its workgroups: [begin / kBlock, blk0 + ceil(count / kBlock))
    float V, rho0, mass, mu, xsph, sigma;  // SPH material (MPSSPHMaterial.h:9-21)
    float dt, gx, gy, gz;
    float bmin[3], bmax[3];
    float cg_factor;
    int faithful;
};

// Scene lookup handed to the fluid kernels by value.
struct Scenes {
    const StepParams* P;        // [S], device
    const uint32_t* blk_scene;  // [ceil(N / kBlock)]: scene of each workgroup of slots
    uint32_t N;                 // fluid slots of the context
    uint32_t S;                 // scenes
};

// Column-major ELL neighbour lists: nbr[k*N + i], k < cnt[i]. fr/cfr: the ref lists into the
// other scenes when the scenes are coupled fluid objects (null otherwise).
struct Nbr {
    const uint32_t* ff;
    const uint32_t* cff;
    const uint32_t* fb;
    const uint32_t* cfb;
    const uint32_t* fr = nullptr;
    const uint32_t* cfr = nullptr;
    // wall objects (null: one object): object of each wall particle, and their number (<= 4)
    const uint32_t* fbo = nullptr;
    uint32_t nbo = 1;
};
constexpr uint32_t kMaxWallObjects = 4;

// Block halos of the fluid-fluid lists (halo.hip): for workgroup b of slots, idx[b*kHaloCap + t]
// (t < n[b]) are the distinct neighbour slots in ascending order and nb16[k*N + i] is the
// position of nff[k*N + i] in its block's list. n[b] > kHaloCap: not staged (use nff).
// Each staged kernel stages the halos up to its own LDS depth and gathers deeper ones from
// global memory. A.p comes in two depths: kHaloShallow (five workgroups per CU) and kHaloCap
// (four); the deep one pays once a few blocks in a thousand have halos over kHaloShallow.
constexpr uint32_t kHaloCap = 1536;
constexpr uint32_t kHaloShallow = 1264;
struct Halo {
    const uint32_t* idx;
    const uint32_t* n;
    const uint16_t* nb16;
};

namespace launch {
// ---- kernel timing
// While a context times its kernels it points `timing` (per host thread) at a Timing; every
// launch below then asks it for an event pair and is dispatched with hipExtLaunchKernel, which
// stamps the pair with the dispatch's own start and end. No marker packets (and none of the
// cache write-back they carry) sit between the kernels, so the durations are those of the
// kernels as they run untimed.
struct Timing {
    virtual void events(hipEvent_t* start, hipEvent_t* stop) = 0;

protected:
    ~Timing() = default;
};
extern thread_local Timing* timing;
// A launch the runtime refused (bad configuration, missing code object): thrown by launch_k,
// turned into MPS_ERR_HIP with the message by the C ABI.
struct LaunchError {
    hipError_t e;
};

// Launch `kernel` on `s` with the arguments converted to its parameter types, as a call would.
template <class... P, class... A>
inline void launch_k(void (*kernel)(P...), dim3 grid, dim3 block, uint32_t shm, hipStream_t s, A&&... args) {
    static_assert(sizeof...(P) == sizeof...(A), "argument count");
    std::tuple<std::decay_t<P>...> t(std::forward<A>(args)...);
    void* ptrs[sizeof...(P) + 1];
    std::apply([&](auto&... a) {
        size_t k = 0;
        ((ptrs[k++] = static_cast<void*>(&a)), ...);
    }, t);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timing) timing->events(&e0, &e1);
    const hipError_t rc = e0 ? hipExtLaunchKernel(reinterpret_cast<const void*>(kernel), grid, block, ptrs, shm, s, e0, e1, 0)
                             : hipLaunchKernel(reinterpret_cast<const void*>(kernel), grid, block, ptrs, shm, s);
    if (rc != hipSuccess) throw LaunchError{rc};  // faults while running surface at the next synchronisation
}

// ---- hash.hip
size_t radix_temp_words(uint32_t n);
int radix_sort(uint32_t*& keys, uint32_t*& vals, uint32_t*& keys_alt, uint32_t*& vals_alt, uint32_t* temp, uint32_t n,
               int key_bits, hipStream_t s);
// Keys carry the scene index above `shift` cell bits (scene / blk_scene may be null for one scene).
void cell_keys(const float4* pos, uint32_t n, const dev::HashGeom& g, int morton, const uint32_t* scene, int shift,
               uint32_t* key, uint32_t* id, hipStream_t s);
void cell_keys_slots(const float4* pos, const uint32_t* perm, uint32_t n, const dev::HashGeom& g, int morton,
                     const uint32_t* blk_scene, int shift, uint32_t* keyref, uint32_t* iota, hipStream_t s);
void cell_morton_slots(const float4* pos, const uint32_t* perm, uint32_t n, const dev::HashGeom& g, const uint32_t* blk_scene,
                       int shift, uint32_t* keyref, uint32_t* iota, hipStream_t s);
// sorted position q -> slot q + delta[key[q] >> shift]: new perm, src (old slot) and key per slot
void place(const uint32_t* key, const uint32_t* id, uint32_t nref, int shift, const uint32_t* delta, const uint32_t* inv,
           uint32_t* perm, uint32_t* src, uint32_t* mkey, hipStream_t s);
void cell_table_m(const uint32_t* mkey, const float4* pos, uint32_t n, const dev::HashGeom& g, int shift, uint32_t C,
                  uint32_t* start, uint32_t* end, uint32_t* lkey, hipStream_t s);
void clear_cells_m(const uint32_t* mkey, const uint32_t* lkey, uint32_t n, uint32_t* start, uint32_t* end, hipStream_t s);
void invert(const uint32_t* perm, uint32_t n, uint32_t* inv, hipStream_t s);
void compose(const uint32_t* perm_new, const uint32_t* inv, uint32_t n, uint32_t* src, hipStream_t s);
void relabel(uint32_t* perm, const uint32_t* map, uint32_t n, hipStream_t s);
void iota(uint32_t* p, uint32_t n, hipStream_t s);
void fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t s);
void clear_cells(const uint32_t* key, uint32_t n, uint32_t* start, uint32_t* end, hipStream_t s);
void cell_table(const uint32_t* key, const uint32_t* id, uint32_t n, uint32_t* start, uint32_t* end, const float4* pos,
                float4* spos, hipStream_t s);
// bstat: scratch of 2 * ceil(N / kBlock) (fluid) or ceil(B / kBlock) (walls) per-workgroup
// count records, folded into ctr after the search
void neighbors_fluid(const float4* pos, const Scenes& sc, uint32_t C, const dev::HashGeom& g, const HashView& fl,
                     const HashView& bd, uint32_t* nff, uint32_t Kff, uint32_t* cff, uint32_t* nfb, uint32_t Kfb, uint32_t* cfb,
                     uint2* bstat, NbrCounters* ctr, hipStream_t s);
// ref lists of coupled scenes (other scenes' fluid, ascending scene order)
void neighbors_ref(const float4* pos, const Scenes& sc, uint32_t C, const dev::HashGeom& g, const HashView& fl, uint32_t* nfr,
                   uint32_t Kfr, uint32_t* cfr, uint2* bstat, NbrCounters* ctr, hipStream_t s);
void neighbors_boundary(const float4* pos, uint32_t B, const uint32_t* scene, uint32_t C, const dev::HashGeom& g,
                        const HashView& bd, uint32_t* nbb, uint32_t Kbb, uint32_t* cbb, uint2* bstat, NbrCounters* ctr,
                        hipStream_t s);
void gather_f4(const uint32_t* perm, uint32_t n, const float4* in, float4* out, hipStream_t s);
void gather_u32(const uint32_t* perm, uint32_t n, const uint32_t* in, uint32_t* out, hipStream_t s);
void gather_f2(const uint32_t* perm, uint32_t n, const float2* in, float2* out, hipStream_t s);
void gather_f1(const uint32_t* perm, uint32_t n, const float* in, float* out, hipStream_t s);

// ---- sph.hip
enum DensityFlags { kDoDensity = 1, kDoAlpha = 2, kDoGravity = 4, kDoPairs = 8 };
void boundary_sum(int cubic, const float4* bpos, uint32_t B, const uint32_t* nbb, uint32_t Kbb, const uint32_t* cbb, float* S,
                  hipStream_t s);
void boundary_volume(const float* S, float* vol, uint32_t B, int faithful, hipStream_t s);
// several wall objects: per-object sums S4[i] = (sum over object 0, 1, 2, 3), W(0) in the own
// object's; volume = 1/(((V_prev + own) + other objects in order))
void boundary_sum_obj(int cubic, const float4* bpos, uint32_t B, const uint32_t* nbb, uint32_t Kbb, const uint32_t* cbb,
                      const uint32_t* obj, float4* S4, hipStream_t s);
void boundary_volume_obj(const float4* S4, const uint32_t* obj, uint32_t nobj, float* vol, uint32_t B, int faithful,
                         hipStream_t s);
// density, alpha, gravity; with kDoPairs also G_ij per neighbour slot
void density_alpha(int cubic, int flags, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, float4* vel,
                   const float4* bpos, const float* bvol, float* rho, float* alpha, float4* color, float* pG_ff,
                   float* pG_fb, hipStream_t s);
// nb.fr set: the coupled-scene variants (ref sums, G_ij evaluated for them with `cubic`)
void pressure_eval(int cubic, int df, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const float4* vel,
                   const float4* bpos, const float* bvol, const float4* bvel, const float* rho, const float* alpha,
                   const float* pG_ff, const float* pG_fb, float* kappa, float* delta, const SolverDev* st, double* partials,
                   hipStream_t s);
// joint: one decision for all scenes (coupled fluid objects)
void fold_pressure(const Scenes& sc, const double* partials, SolverDev* st, uint32_t* active, double tol, uint32_t lmin,
                   int joint, hipStream_t s);
void pressure_apply(int cubic, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, float4* vel, const float4* bpos,
                    const float* bvol, const float* pG_ff, const float* pG_fb, const float* kappa, const SolverDev* st,
                    hipStream_t s);
void st_factor(const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const float4* vel, const float* pG_ff, float* fst,
               float4* cgx, hipStream_t s);
// pair coefficients of the implicit operator: a_ij into ca and/or (a_ij, b_ij) into cab (either may be null)
void cg_setup(int cubic, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const float4* vel, const float* rho,
              const float* fst, float* ca, float2* cab, float4* cgb, float* irho, hipStream_t s);
// The adjoint's transposed solve keeps p in 32-byte records with the position:
// rec[2i] = (x, y, z, p.x), rec[2i+1] = (p.y, p.z, 0, 0)
void cg_pack(uint32_t N, const float4* pos, const float4* x0, float4* rec, hipStream_t s);
// The forward solve's vectors are packed triples (dev::f3).
// mode 0: A x0 with f = 0 (x0 from the float4 array x0), then r = b - A x0, pout = r, x = x0, partial r.r;
// mode 1: p = r + beta pprev (r in rr, beta from the record), A p with f into Ap, p into pout,
// partial p.Ap. Halo records are staged in LDS (kHaloCap).
// The implicit operator's coefficients: streamed (a_ij, b_ij) pairs (cab), or a_ij (ca) with
// b_ij recomputed from positions, masses (vel.w) and 1/rho (irho) of the CG set-up.
struct CgOperator {
    int recompute = 0;
    int cubic = 1;
    int deep = 0;  // the streamed operator stages halos up to kHaloCap (else kHaloShallow)
    const float2* cab = nullptr;
    const float* ca = nullptr;
    const float4* vel = nullptr;
    const float* rho = nullptr;
    const float* irho = nullptr;
};
void cg_apply(int mode, const CgOperator& op, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos,
              const float4* x0, const dev::f3* rr, const dev::f3* pprev, dev::f3* pout, dev::f3* Ap, const float4* b,
              dev::f3* r, dev::f3* x, const SolverDev* st, double* partials, hipStream_t s);
// A.p (mode 1) and the update, each with its fold (k_fold_cg 1 / 2) done by the scene's last
// workgroup to finish (the solver record's counter counts arrivals): two launches per CG
// iteration instead of four, with the same bits. arrive: one counter per workgroup (zero, and
// zero again after every launch).
void cg_apply_fold(const CgOperator& op, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, const dev::f3* rr,
                   const dev::f3* pprev, dev::f3* pout, dev::f3* Ap, SolverDev* st, double* partials, uint32_t* arrive,
                   hipStream_t s);
void cg_update3_fold(const Scenes& sc, dev::f3* x, dev::f3* r, const dev::f3* p, const dev::f3* Ap, SolverDev* st,
                     double* partials, uint32_t* active, double tol, uint32_t* arrive, hipStream_t s);
// Workgroups of the persistent CG kernel that fit the device at once (0 if unknown).
// coherent: the barrier variant (0: agent-scope fences; 1: device-coherent accesses of the data
// that crosses workgroups, no cache write-back or invalidation)
uint32_t cg_persistent_capacity(const CgOperator& op, int coherent);
// CG iterations 1..cap in one cooperative launch of one workgroup per 256-slot block (all must be
// co-resident: nblk <= cg_persistent_capacity). r, the two p buffers (p_0 = r in pbuf0), A p and
// x as in cg_apply / cg_update3; partA/partB: two partial-sum buffers of nblk doubles; gs: the
// barrier words; spin: polls before a barrier wait gives up (gs->gave_up is then set).
void cg_persistent(const CgOperator& op, int coherent, const Scenes& sc, const Nbr& nb, const Halo& hl, const float4* pos, dev::f3* r,
                   dev::f3* pbuf0, dev::f3* pbuf1, dev::f3* Ap, dev::f3* x, SolverDev* st, uint32_t* active, double* partA,
                   double* partB, GridSync* gs, uint32_t cap, double tol, uint32_t spin, hipStream_t s);
// ---- halo.hip
// stats (3 words, accumulated): largest staged halo, blocks over kHaloCap, blocks with particles
// limit: largest halo staged (0 or > kHaloCap: kHaloCap)
void halo_build(const Scenes& sc, const Nbr& nb, uint32_t nblk, uint32_t limit, uint32_t* hidx, uint32_t* hn, uint16_t* n16,
                uint32_t* stats, hipStream_t s);
// mode 0: initial r.r + test; 1: alpha; 2: new r.r, beta + test
void fold_cg(int mode, const Scenes& sc, const double* partials, SolverDev* st, uint32_t* active, double tol, hipStream_t s);
// x += alpha p, r -= alpha Ap, partial r.r (p from the records; the adjoint's solve)
void cg_update(const Scenes& sc, float4* x, float4* r, const float4* prec, const float4* Ap, const SolverDev* st,
               double* partials, hipStream_t s);
// the same on packed triples (the forward solve), and its copy of x into the velocities
void cg_update3(const Scenes& sc, dev::f3* x, dev::f3* r, const dev::f3* p, const dev::f3* Ap, const SolverDev* st,
                double* partials, hipStream_t s);
void copy_vel_out3(const Scenes& sc, const dev::f3* x, float4* vel, hipStream_t s);
// p = r + beta p in the records (the adjoint's solve; the forward forms p inside cg_apply)
void cg_direction(const Scenes& sc, const float4* r, float4* prec, const SolverDev* st, hipStream_t s);
void copy_vel_out(const Scenes& sc, const float4* x, float4* vel, hipStream_t s);
void jacobi_candidate(int cubic, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel,
                      const float* rho, const float4* u, float4* out, hipStream_t s);
void jacobi_relax(uint32_t N, const float4* cand, const float4* cur, const float4* prev, float ur, const float* w3,
                  float4* o0, float4* o1, float4* o2, hipStream_t s);
// sums |u - a sum(...) - v0| for up to three candidates into err3 (double, deterministic)
void jacobi_error(int cubic, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel,
                  const float* rho, const float4* u0, const float4* u1, const float4* u2, int ncand, double* partials,
                  double* err3, hipStream_t s);
void advect(const Scenes& sc, float4* pos, float4* vel, uint32_t* nonfinite, hipStream_t s);
// resets the S solver records st[0..S) and sets *active = S
void solver_reset(SolverDev* st, uint32_t S, uint32_t* active, hipStream_t s);
void calib_copy(const float4* a, float4* b, size_t n, hipStream_t s);
// STREAM-style copy (one 16-byte vector per thread, non-temporal): the in-run copy rate
void stream_copy(const float4* a, float4* b, size_t n, hipStream_t s);
void zero_words(uint32_t* p, size_t nwords, hipStream_t s);
void pack_readback(uint32_t n, const float4* pos, const float4* color, const uint32_t* perm, float* xyz, float* rad,
                   float* rgba, hipStream_t s);
void pack_readback_f64(uint32_t n, const float4* pos, const float4* color, const uint32_t* perm, double* xyz, double* rad,
                       float* rgba, hipStream_t s);

// ---- adjoint.hip (reverse mode of one step with frozen control flow)
// the kernel set (cubic spline, or poly6 density with the spiky gradient) and whether every
// scene's fluid radii are equal (then a pair's G serves both ends)
struct AdjKind {
    int cubic, eq;
};
void adj_advect(const Scenes& sc, const float4* pos, const float4* vin, const float4* ax, float4* av, hipStream_t s);
// aux = (kappa, e, mask, rho branch) recomputed from the velocity an iteration started from
void adj_peval(int df, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel, const float4* bpos,
               const float* bvol, const float4* bvel, const float* rho, const float* alpha, const float* pG_ff,
               const float* pG_fb, float4* aux, hipStream_t s);
void adj_kbar(AdjKind k, int df, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* bpos, const float* bvol,
              const float* pG_ff, const float* pG_fb, const float4* aux, const float* alpha, const float4* av,
              const uint32_t* L, uint32_t it, float* dbar, float4* ax, float* abar, float* rbar, hipStream_t s);
void adj_dbar(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vel, const float4* bpos,
              const float* bvol, const float4* bvel, const float* pG_ff, const float* pG_fb, const float* dbar, const uint32_t* L, uint32_t it,
              float4* av, float4* ax, hipStream_t s);
void adj_rev_index(const Scenes& sc, const Nbr& nb, uint32_t* rev, uint32_t* bad, hipStream_t s);
void adj_transpose_coeffs(const Scenes& sc, const Nbr& nb, const uint32_t* rev, const float2* cab, float2* cabT, hipStream_t s);
void adj_cg_init(const Scenes& sc, const float4* pos, const float4* b, float4* x, float4* r, float4* rec, double* partials,
                 hipStream_t s);
void adj_cg_applyT(const Scenes& sc, const Nbr& nb, const float2* cab, const float2* cabT, const float4* rec, float4* Ap,
                   const SolverDev* st, double* partials, hipStream_t s);
// partials: 2 * ceil(N / kBlock) doubles (mubar, sigmabar per workgroup)
void adj_cg_grad(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vin, const float4* vstar, const float* rho,
                 const float* fst, const float4* lam, float4* avin, float4* ax, float* rbar, float* fbar, double* partials,
                 hipStream_t s);
void adj_fst(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* vin, const float* pG_ff, const float* fbar,
             float4* avin, float4* ax, hipStream_t s);
void adj_fold2(const Scenes& sc, const double* partials, double* acc0, double* acc1, hipStream_t s);
void adj_gvec(const Scenes& sc, const Nbr& nb, const float4* pos, const float4* bpos, const float* bvol, const float* pG_ff,
              const float* pG_fb, const float* abar, float4* gv, hipStream_t s);
void adj_density(AdjKind k, const Scenes& sc, const Nbr& nb, const float4* pos, const float4* bpos, const float* bvol, const float* pG_ff,
                 const float* pG_fb, const float* rbar, const float4* gv, float4* ax, hipStream_t s);
// also carries the transposed solve's solution (al, may be null) back, as the next warm start
void adj_unpermute(uint32_t n, const uint32_t* src, const float4* ax, const float4* av, const float4* al, float4* ax_out,
                   float4* av_out, float4* al_out, hipStream_t s);
// warm start of A_f^T lambda = b from x0 (packed in recA with the positions): partials of r.r and b.b
void adj_cg_init_warm(const Scenes& sc, const Nbr& nb, const float2* cab, const float2* cabT, const float4* recA,
                      const float4* b, float4* x, float4* r, float4* recB, double* partials, hipStream_t s);
void adj_fold_start(const Scenes& sc, const double* partials, SolverDev* st, uint32_t* active, double tol, hipStream_t s);
// adx[s] += gx[perm[s]], adv[s] += gv[perm[s]] (gx, gv: n_ref x 3 floats in reference order; gv may be
// null): the cotangent of an observed state, added in the slot order that state was stored in
void adj_add_cotangent(uint32_t n, const uint32_t* perm, const float* gx, const float* gv, float4* adx, float4* adv,
                       hipStream_t s);

// ---- sampling.hip
// Face ownership bits of a triangle mesh (bits 0-2 owned vertices, 3-5 owned edges) on the GPU;
// scratch: vmin (nv), ekey/emin (owner_table_slots(nt) entries each)
size_t owner_table_slots(uint32_t nt);
void owner_build(const uint32_t* F, uint32_t nt, uint32_t nv, uint32_t* vmin, unsigned long long* ekey, uint32_t* emin,
                 size_t slots, uint32_t* R, hipStream_t s);
void sample_count(const double* V, const uint32_t* F, const uint32_t* R, uint32_t nt, double radius, uint32_t* counts,
                  uint32_t* shorts, hipStream_t s);
void sample_generate(const double* V, const uint32_t* F, const uint32_t* R, uint32_t nt, double radius, const uint8_t* gen,
                     const uint32_t* offs, double* pos, uint32_t* face, uint32_t* bad, hipStream_t s);
void sample_finish(const double* V, const uint32_t* F, uint32_t nt, const double* vmass, uint32_t n, const double* pos,
                   const uint32_t* face, double radius, int keep_vel, float4* bpos, float4* bvel, float2* bbcc,
                   uint32_t* bface, float4* bcolor, hipStream_t s);
// re-sampling: slots of faces not generated anew take their particle from the previous arrays
void sample_copy(uint32_t nt, const uint32_t* curr, const uint32_t* prev, const uint8_t* gen, uint32_t n,
                 const float4* prev_bpos, const float2* prev_bbcc, const uint32_t* prev_bface, const float4* prev_bcolor,
                 const float4* prev_bvel, float4* bpos, float2* bbcc, uint32_t* bface, float4* bcolor, float4* bvel,
                 hipStream_t s);
}  // namespace launch
}  // namespace mps
